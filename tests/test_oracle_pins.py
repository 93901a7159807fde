"""Closed-form pins of the oracle's Stage A and Monte Carlo move, written without the oracle's code.

Stage A (src/adock/mckernel.py:451-521): per (complex, conformer, rotation) row, the minimum over the translation
lattice of sum_a min(nearest-grid energy + out-of-box penalty, cap), and the lowest lattice index reaching it.
Checked on grids whose nearest-point values are known in closed form (constant, linear in the point index), with
atoms placed at 0.4 / 0.6 of a cell so that round-to-nearest and floor disagree, distinct coefficients per axis (an
index transposition changes the sums), a cap below single-atom energies (a cap applied to the sum instead of per atom
changes the result), atoms pushed past the box by a known distance (penalty pen * d^2), duplicated lattice
translations (ties go to the lowest index), and a brute-force numpy argmin on a tiny random lattice.

MC move (src/adock/mckernel.py:379-398): on a zero grid with a ligand that has no intra pairs, BFGS cannot move the
pose (zero gradient) and every move is accepted (equal energies, u < exp(0) = 1), so after one step the chain's state
is exactly the mutated start. The expected mutation is recomputed here from the counter RNG (SplitMix64 written in
Python, checked against its published vectors): entity from slot 0 over 2 + nt choices, unit-ball point from slots
10 + 3k (first of up to 32 tries inside the ball), translation by amp_pos * v, rotation quat(amp_rot / max(rg, 1) * v)
* q, torsion set to (2 u_1 - 1) pi. On a terraced grid (flat steps one cell apart) a one-atom ligand's energy after a
translation is known without the oracle, which pins the Metropolis test against slot 2."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from adock_amd.config import search_prm, sf_params
from adock_amd.ligand import LigandPrep, pack_ligands
from helpers import CFG, IDX

GOLD = Path(__file__).resolve().parent / "golden"
M64 = (1 << 64) - 1


# ---------------------------------------------------------------------------------------------------- counter RNG
def smix_py(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def urand_py(seed, chain, step, slot):
    ctr = ((chain & 0xFFFFFFFF) << 32) ^ ((step << 8) & M64) ^ (slot & M64)
    h = smix_py((seed ^ smix_py(ctr)) & M64)
    return np.float32((h >> 40) * (1.0 / 16777216.0))


def test_python_splitmix_matches_published_vectors():
    g = json.loads((GOLD / "splitmix64.json").read_text())
    for x, want in g["smix"]:
        assert smix_py(int(x, 16)) == int(want, 16)
    for c, s, k in [(0, 0, 0), (5, 17, 2), (1 << 20, 999, 40), (3, 0, 10)]:
        assert float(urand_py(12345, c, s, k)) == oracle.urand(12345, c, s, k)


def ball_py(seed, chain, step, slot=10):
    for k in range(32):
        v = [np.float32(2.0) * urand_py(seed, chain, step, slot + 3 * k + i) - np.float32(1.0) for i in range(3)]
        if float(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) < 1.0:
            return np.array(v, np.float64)
    return np.zeros(3)


def quat_mul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return np.array([aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw])


def rotvec_quat(r):
    ang = float(np.linalg.norm(r))
    if ang <= 1e-6:
        return np.array([1.0, 0.5 * r[0], 0.5 * r[1], 0.5 * r[2]])
    return np.concatenate([[math.cos(ang / 2)], np.sin(ang / 2) / ang * np.asarray(r)])


# ---------------------------------------------------------------------------------------------------- tiny problems
SPACING = 0.375
PEN = 10.0


def _sf():
    sfp = sf_params(CFG, "vina")
    return dict(prm=search_prm(sfp), rad=np.asarray(sfp["radius"], np.float32),
                hyd=np.asarray(sfp["hydrophobic"], np.int32), don=np.asarray(sfp["donor"], np.int32),
                acc=np.asarray(sfp["acceptor"], np.int32))


def _grid(G, origin, lo, hi):
    G = np.asarray(G, np.float32)
    return dict(flat=G.reshape(-1), origin=np.asarray(origin, np.float32)[None], lo=np.asarray(lo, np.float32)[None],
                hi=np.asarray(hi, np.float32)[None], rec_idx=np.zeros(1, np.int32), offs=np.zeros(1, np.int64),
                dims=np.array([G.shape[1:]], np.int64), spacing=SPACING, penalty=PEN)


def _problem(preps, G, origin, lo, hi, rg=None):
    T = G.shape[0]
    slot = -np.ones((1, len(CFG["_xs_names"])), np.int64)
    types = sorted({int(t) for p in preps for t in p.types[p.types >= 0]})
    assert len(types) <= T
    slot[0, types] = np.arange(len(types))
    AM, KM, lig = pack_ligands(preps, slot, np.zeros(len(preps), np.int64))
    if rg is not None:
        lig["rg"][:] = rg
    return oracle.OracleProblem(AM, KM, lig, _grid(G, origin, lo, hi), _sf()), lig


def star_ligand(arms, bond=1.5, t="C_P"):
    """A centre atom bonded to `arms` atoms, every bond rotatable: `arms` torsions and no intra pairs (every pair of
    atoms is at most 2 bonds apart)."""
    dirs = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)[:arms]
    xyz = np.concatenate([[[0.0, 0.0, 0.0]], bond * dirs])
    bonds = [(0, i + 1) for i in range(arms)]
    return LigandPrep(xyz, [IDX[t]] * (arms + 1), bonds, bonds)


def one_atom(t="C_P"):
    return LigandPrep(np.zeros((1, 3)), [IDX[t]], [], [])


# ---------------------------------------------------------------------------------------------------- Stage A
def _stage_a(op, conf, quats, trans, ntrans, cap):
    be, bi = op.stage_a(conf, quats, trans, ntrans, cap)
    return np.asarray(be), np.asarray(bi)


def _lattice(T):
    T = np.asarray(T, np.float32)
    return T[None], np.array([len(T)], np.int32)


def test_stage_a_constant_grid_caps_and_ties():
    """Constant grid value c per type, every translation inside the box: e = sum_a min(c, cap) for every translation,
    so the argmin is translation 0 (lowest index); with c above the cap each atom contributes the cap."""
    n = 24
    prep = star_ligand(4, t="C_P")     # 5 atoms, one type
    na = 5
    origin = np.zeros(3)
    lo, hi = origin, origin + (n - 1) * SPACING
    conf = np.zeros((1, 1, 64, 3), np.float32)
    conf[0, 0, :na] = prep.ref[:na]
    quats = np.array([[1, 0, 0, 0], [0.5, 0.5, 0.5, 0.5]], np.float32)
    trans, nT = _lattice([[4.3, 4.3, 4.3], [4.0, 4.6, 4.3], [4.5, 4.1, 3.9]])
    for c, cap, want in [(-0.7, 2.0, -0.7 * na), (3.0, 2.0, 2.0 * na), (1.5, 1.4, 1.4 * na)]:
        G = np.full((1, n, n, n), c, np.float32)
        op, _ = _problem([prep], G, origin, lo, hi)
        be, bi = _stage_a(op, conf, quats, trans, nT, cap)
        np.testing.assert_allclose(be, want, rtol=1e-6, atol=1e-6)
        assert (bi == 0).all()


def test_stage_a_linear_grid_nearest_point_penalty_and_cap():
    """G[i, j, k] = a i + b j + c k (a, b, c distinct) with atoms at 0.4 and 0.6 of a cell past grid points: the
    nearest point of an atom at (i + f) h is i for f = 0.4 and i + 1 for f = 0.6 (floor would give i both times).
    One translation puts an atom 0.95 A past the upper x face (index clamped to the last point, penalty pen * 0.95^2).
    The cap applies per atom: with cap 0.5 that row is min(e_0, 0.5) + min(e_1 + 9.025, 0.5)."""
    n = 20
    a, b, c = 0.03, -0.05, 0.011
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    G = (a * ii + b * jj + c * kk).astype(np.float32)[None]
    origin = np.array([1.0, -2.0, 0.5])
    lo, hi = origin, origin + (n - 1) * SPACING
    # a two-atom rigid ligand; its atoms sit at these offsets (in cells) from the translation point
    prep = LigandPrep(np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0]]), [IDX["C_P"]] * 2, [(0, 1)], [])
    off = np.array([[0.4, 1.6, 2.4], [3.6, 0.4, 5.6]])
    conf = np.zeros((1, 1, 64, 3), np.float32)
    conf[0, 0, :2] = off * SPACING
    quats = np.array([[1, 0, 0, 0]], np.float32)
    cells = np.array([[5, 5, 5], [2, 9, 3], [10, 1, 7], [n - 1 - 3, 4, 4]], np.float64)
    trans = origin + cells * SPACING
    trans[3, 0] += 0.5 + 0.6 * SPACING

    def closed_form(t, cap):
        e = 0.0
        for at in range(2):
            x = origin + off[at] * SPACING + (trans[t] - origin)
            idx = np.clip(np.floor((x - origin) / SPACING + 0.5).astype(int), 0, n - 1)
            ea = a * idx[0] + b * idx[1] + c * idx[2]
            ea += PEN * float((np.maximum(x - hi, 0) ** 2 + np.maximum(lo - x, 0) ** 2).sum())
            e += min(ea, cap)
        return e

    # by hand for translation 0: atom 0 at cells (5.4, 6.6, 7.4) -> point (5, 7, 7); atom 1 at (8.6, 5.4, 10.6) ->
    # (9, 5, 11); no penalty
    assert abs(closed_form(0, 100.0) - ((a * 5 + b * 7 + c * 7) + (a * 9 + b * 5 + c * 11))) < 1e-12
    # translation 3 (shifted by 0.5 A + 0.6 cells along x): atom 0 at cells (18.33, 5.6, 6.4) -> point (18, 6, 6);
    # atom 1 at (21.53, 4.4, 9.6), i.e. x = hi + 0.95 A, clamped to point (19, 4, 10), penalty 10 * 0.95^2
    e1 = a * 19 + b * 4 + c * 10 + PEN * 0.95 ** 2
    e0 = a * 18 + b * 6 + c * 6
    assert abs(closed_form(3, 100.0) - (e0 + e1)) < 1e-9
    assert abs(closed_form(3, 0.5) - (min(e0, 0.5) + 0.5)) < 1e-12
    op, _ = _problem([prep], G, origin, lo, hi)
    T = trans.astype(np.float32)[None]
    for cap in (0.5, 100.0):
        want = [closed_form(t, cap) for t in range(len(trans))]
        be, bi = _stage_a(op, conf, quats, T, np.array([len(trans)], np.int32), cap)
        assert int(bi[0, 0]) == int(np.argmin(want))
        np.testing.assert_allclose(be[0, 0], min(want), rtol=1e-6, atol=1e-7)
        # each translation alone (one-point lattices) reproduces its closed form
        for t in range(len(trans)):
            be1, _ = _stage_a(op, conf, quats, T[:, t:t + 1], np.array([1], np.int32), cap)
            np.testing.assert_allclose(be1[0, 0], want[t], rtol=1e-6, atol=1e-7)


def test_stage_a_tie_goes_to_lowest_index_and_brute_force():
    """Duplicated lattice translations tie; the reported index is the lowest one. A brute-force numpy argmin over a
    random grid, rotations and lattice agrees with the oracle row for row."""
    rng = np.random.default_rng(4)
    n = 16
    G = rng.uniform(-1, 1, (2, n, n, n)).astype(np.float32)
    origin = np.zeros(3)
    lo, hi = origin + 0.3, origin + (n - 1) * SPACING - 0.3
    prep = LigandPrep(np.array([[0.0, 0.0, 0.0], [1.4, 0.2, 0.0], [2.0, 1.3, 0.1]]),
                      [IDX["C_P"], IDX["N_A"], IDX["C_P"]], [(0, 1), (1, 2)], [])
    op, lig = _problem([prep], G, origin, lo, hi)
    na = 3
    conf = np.zeros((1, 2, 64, 3), np.float32)
    conf[0, 0, :na] = prep.ref
    conf[0, 1, :na] = prep.ref[::-1] * 0.9
    q = rng.standard_normal((5, 4))
    quats = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
    trans = rng.uniform(0.5, (n - 1) * SPACING - 0.5, (30, 3)).astype(np.float32)
    trans[17] = trans[4]      # duplicate: ties with index 4
    T, nT = _lattice(trans)
    cap = 0.6
    be, bi = _stage_a(op, conf, quats, T, nT, cap)
    gslot = lig["gslot"][0, :na]
    # the kernels take the cap and the box in float32
    cap, lo, hi = float(np.float32(cap)), lo.astype(np.float32).astype(np.float64), hi.astype(np.float32).astype(
        np.float64)
    for c in range(2):
        for qi in range(5):
            w, x, y, z = quats[qi].astype(np.float64) / np.linalg.norm(quats[qi].astype(np.float64))
            Rm = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                           [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                           [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
            P = conf[0, c, :na].astype(np.float64) @ Rm.T
            E = np.zeros(len(trans))
            for t in range(len(trans)):
                X = P + trans[t]
                idx = np.clip(np.rint((X - origin) / SPACING).astype(int), 0, n - 1)
                ea = G[gslot, idx[:, 0], idx[:, 1], idx[:, 2]].astype(np.float64)
                ea += PEN * ((np.maximum(lo - X, 0) ** 2 + np.maximum(X - hi, 0) ** 2).sum(1))
                E[t] = np.minimum(ea, cap).sum()
            row = c * 5 + qi
            assert int(bi[0, row]) == int(np.argmin(E)), (c, qi)
            np.testing.assert_allclose(be[0, row], E.min(), rtol=1e-9, atol=1e-9)
            if np.argmin(E) == 4:
                assert bi[0, row] == 4 and E[17] == E[4]
    # the duplicate translation ties exactly, so a lattice holding it twice reports the first copy
    one = np.stack([trans[4], trans[4]])[None]
    _, bi2 = _stage_a(op, conf, quats, one, np.array([2], np.int32), cap)
    assert (bi2 == 0).all()


# ---------------------------------------------------------------------------------------------------- MC move
def _expected_move(x0, seed, cid, sg, nt, rg, amp_pos, amp_rot):
    x = np.array(x0, np.float64)
    nm = 2 + nt
    which = min(int(np.float32(urand_py(seed, cid, sg, 0) * np.float32(nm))), nm - 1)
    if which == 0:
        x[:3] += amp_pos * ball_py(seed, cid, sg)
    elif which == 1:
        v = ball_py(seed, cid, sg)
        q = quat_mul(rotvec_quat(amp_rot / max(rg, 1.0) * v), x[3:7])
        x[3:7] = q / np.linalg.norm(q)
    else:
        x[7 + which - 2] = (2.0 * float(urand_py(seed, cid, sg, 1)) - 1.0) * 3.14159265
    x[3:7] /= np.linalg.norm(x[3:7])
    return x, which


@pytest.mark.parametrize("rg", [0.6, 2.5])
@pytest.mark.parametrize("f64", [True, False])
def test_mc_move_on_flat_landscape(rg, f64):
    """Zero grid, star ligand (4 torsions, no pairs), poses deep inside the box: after one MC step every chain holds
    exactly the mutated start state (BFGS has zero gradient, the move is always accepted)."""
    n = 120
    G = np.zeros((1, n, n, n), np.float32)
    origin = np.zeros(3)
    lo, hi = origin, origin + (n - 1) * SPACING
    prep = star_ligand(4)
    op, lig = _problem([prep], G, origin, lo, hi, rg=rg)
    KM = op.KM
    N = 400
    rng = np.random.default_rng(11)
    st = np.zeros((N, 7 + KM), np.float32)
    st[:, :3] = (hi / 2 + rng.uniform(-3, 3, (N, 3))).astype(np.float32)
    q = rng.standard_normal((N, 4))
    st[:, 3:7] = q / np.linalg.norm(q, axis=1, keepdims=True)
    st[:, 7:11] = rng.uniform(-np.pi, np.pi, (N, 4))
    cur, best = st.copy(), st.copy()
    ce, be = np.zeros(N, np.float32), np.zeros(N, np.float32)
    seed, sg0 = 987654321, 37
    cids = 1000 + 3 * np.arange(N)
    op.run_mc(np.zeros(N), cids, cur, ce, best, be, False, 1, 1, 1.2, 2.0, 2.0, seed, sg0, f64=f64)
    seen = set()
    for i in range(N):
        want, which = _expected_move(st[i, :11], seed, int(cids[i]), sg0, 4, rg, 2.0, 2.0)
        seen.add(min(which, 2))
        np.testing.assert_allclose(cur[i, :11], want, rtol=0, atol=2e-5, err_msg=f"chain {i} which {which}")
        assert ce[i] == 0.0
    assert seen == {0, 1, 2}


def test_mc_metropolis_on_terraced_grid():
    """One-atom ligand on a grid that is constant on slabs of 4 x-cells, rising by 1.5 kcal/mol per slab: an atom
    inside a slab has zero gradient (BFGS leaves it), so after one step the chain either holds the translated atom at
    the new slab's energy (accepted: lower energy, or u_2 < exp(-dE / T)) or keeps its start. Checked on every chain
    whose start and translated positions lie strictly inside slabs."""
    nx, n = 64, 12
    step = 1.5
    i = np.arange(nx)
    slab = (i // 4).astype(np.float64)
    G = np.broadcast_to((step * slab)[:, None, None], (nx, n, n)).astype(np.float32)[None].copy()
    origin = np.zeros(3)
    lo, hi = origin - 50.0, origin + 50.0           # no penalty anywhere near
    prep = one_atom()
    op, lig = _problem([prep], G, origin, lo, hi)
    KM = op.KM
    N = 3000
    rng = np.random.default_rng(5)
    st = np.zeros((N, 7 + KM), np.float32)
    st[:, 0] = rng.uniform(6 * SPACING, (nx - 7) * SPACING, N)
    st[:, 1:3] = (n - 1) * SPACING / 2
    st[:, 3] = 1.0
    temp, seed = 1.2, 2024

    def slab_energy(x):
        f = x / SPACING
        i0 = int(np.floor(f))
        if i0 % 4 == 3 or min(f - i0, i0 + 1 - f) < 1e-3 or i0 < 0 or i0 >= nx - 1:
            return None          # a transition cell (non-zero gradient), or too close to a cell face
        return step * (i0 // 4)

    ec0 = np.array([slab_energy(float(x)) if slab_energy(float(x)) is not None else np.nan for x in st[:, 0]],
                   np.float32)
    cur, best = st.copy(), st.copy()
    ce, be = ec0.copy(), ec0.copy()
    op.run_mc(np.zeros(N), np.arange(N), cur, ce, best, be, False, 1, 1, temp, 2.0, 2.0, seed, 0, f64=True)
    checked = accepted = rejected = 0
    for c in range(N):
        if np.isnan(ec0[c]):
            continue
        want, which = _expected_move(st[c, :8], seed, c, 0, 0, 1.0, 2.0, 2.0)
        if which != 0:
            continue
        e_new = slab_energy(float(want[0]))
        if e_new is None:        # (the energy does not depend on y, z: those axes are constant, clamped or not)
            continue
        acc = e_new < ec0[c] or float(urand_py(seed, c, 0, 2)) < math.exp((float(ec0[c]) - e_new) / temp)
        checked += 1
        if acc:
            accepted += 1
            np.testing.assert_allclose(cur[c, :3], want[:3], atol=2e-5)
            assert abs(float(ce[c]) - e_new) < 1e-6
        else:
            rejected += 1
            np.testing.assert_allclose(cur[c, :3], st[c, :3], atol=0)
            assert float(ce[c]) == float(ec0[c])
    assert checked > 500 and accepted > 100 and rejected > 100, (checked, accepted, rejected)
