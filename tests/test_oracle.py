"""Pins the CPU oracle before it is trusted as the parity checker: published SplitMix64 vectors, the reference's
two-atom known-answer table, brute-force grid sums, trilinear exactness on linear fields, finite-difference gradients,
rigid-body invariances, closed-form FK, and BFGS / MC properties."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from adock_amd import synth
from adock_amd.config import sf_params
from helpers import (CFG, IDX, PENALTY, SPACING, analytic_pair, cpu_problem, oracle_grid, random_receptor, states,
                     sf_tables)

GOLD = Path(__file__).resolve().parent / "golden"


def test_splitmix64_published_vectors():
    g = json.loads((GOLD / "splitmix64.json").read_text())
    for x, want in g["smix"]:
        assert oracle.smix(int(x, 16)) == int(want, 16)


def f32_fma(a, b, c):
    """float32 fused multiply-add: a*b is exact in float64 for float32 operands, one rounding to float32 at the end."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def test_urand_is_24_bit_uniform():
    u = np.array([oracle.urand(12345, c, s, k) for c in range(4) for s in range(50) for k in range(5)])
    assert (u >= 0).all() and (u < 1).all()
    assert np.all(np.floor(u * 2 ** 24) == u * 2 ** 24)
    assert abs(u.mean() - 0.5) < 0.05
    # the (chain, step, slot) counter is hashed: neighbouring counters are uncorrelated
    assert abs(np.corrcoef(u[:-1], u[1:])[0, 1]) < 0.1


PAIRS = [("C_H", "C_H"), ("C_H", "C_P"), ("N_D", "O_A"), ("O_DA", "O_DA"), ("N_P", "O_A"), ("Cl_H", "C_H"),
         ("Met_D", "O_A"), ("S_P", "N_DA")]
RS = [1.0, 2.0, 2.9, 3.3, 3.6, 4.0, 4.5, 5.5, 7.99, 8.01]


@pytest.mark.parametrize("sf", ["vina", "vinardo"])
def test_two_atom_known_answers(sf):
    """One receptor atom on the grid's x axis: the grid value at each point is the two-atom energy
    (the reference's table tests/test_scoring_grid.py:16-51)."""
    sfp = sf_params(CFG, sf)
    gold = json.loads((GOLD / "two_atom_terms.json").read_text())[sf]
    for a, b in PAIRS:
        for r in RS:
            rec = {"rec_xyz": np.array([[float(r), 0.0, 0.0]], np.float32), "rec_types": np.array([IDX[b]], np.int32)}
            G = oracle_grid(rec, sfp, np.zeros(3), (1, 1, 1), [IDX[a]])
            want = gold[f"{a}-{b}"][str(r)]
            # the oracle takes the float32 parameter vector: weights carry float32 rounding (<= 6e-8 relative)
            assert abs(G[0, 0, 0, 0] - want) <= 2e-7 * abs(want) + 1e-8, (a, b, r)
            assert abs(want - analytic_pair(sf, float(np.float32(r)), IDX[a], IDX[b])) <= 1e-12 * abs(want) + 1e-15


def test_contact_values_closed_form():
    """d = 0 (atoms in contact): gauss1 = 1, gauss2 = exp(-(3/2)^2), no repulsion, hydrophobic 1 (C_H-C_H)."""
    sfp = sf_params(CFG, "vina")
    r = np.float32(3.8)   # R(C_H) + R(C_H)
    rec = {"rec_xyz": np.array([[float(r), 0, 0]], np.float32), "rec_types": np.array([0], np.int32)}
    G = oracle_grid(rec, sfp, np.zeros(3), (1, 1, 1), [0], per_term=True)
    d = float(r) - 2 * float(np.float32(1.9))
    assert G[0, 0, 0, 0, 0] == pytest.approx(math.exp(-(d / 0.5) ** 2), rel=1e-12)
    assert G[1, 0, 0, 0, 0] == pytest.approx(math.exp(-((d - 3.0) / 2.0) ** 2), rel=1e-12)
    assert G[3, 0, 0, 0, 0] == 1.0 and G[4, 0, 0, 0, 0] == 0.0


@pytest.mark.parametrize("sf", ["vina", "vinardo"])
def test_grid_points_vs_brute_force(sf):
    """Grid values = direct FP64 sum over all receptor atoms (reference recipe: 300 atoms, seed 0, origin (3,3,3),
    n 21, 200 random points, seed 1); checked to 1e-9 relative since both sides are FP64."""
    sfp = sf_params(CFG, sf)
    xyz, types = random_receptor()
    origin, n = np.array([3.0, 3.0, 3.0]), 21
    lt = [IDX[t] for t in ("C_H", "N_DA", "O_A", "F_H")]
    G = oracle_grid({"rec_xyz": xyz, "rec_types": types}, sfp, origin, (n, n, n), lt)
    ijk = np.random.default_rng(1).integers(0, n, (200, 3))
    for q, t in enumerate(lt):
        for i, j, k in ijk:
            p = (origin + np.array([i, j, k]) * SPACING).astype(np.float32)   # exact: multiples of 1/8
            dv = xyz - p
            r2 = f32_fma(dv[:, 0], dv[:, 0], f32_fma(dv[:, 1], dv[:, 1], dv[:, 2] * dv[:, 2]))
            r = np.linalg.norm(dv.astype(np.float64), axis=1)
            terms = [analytic_pair(sf, rr, t, int(tb)) for rr, tb, ok in zip(r, types, r2 < np.float32(64)) if ok]
            direct = sum(terms)
            # float32 weights in the oracle's parameter vector: <= 6e-8 relative per term
            assert abs(G[q, i, j, k] - direct) <= 2e-7 * sum(abs(x) for x in terms) + 1e-9, (t, i, j, k)


def test_grid_invariances():
    """Atom order does not change the FP64 grid; a type subset equals the slices of the full grid; 5 raw channels
    weighted by the term weights equal the weighted grid."""
    sfp = sf_params(CFG, "vina")
    xyz, types = random_receptor(120, seed=3, extent=10.0)
    origin, n = np.array([1.0, 1.5, 2.0]), (9, 11, 13)
    full = oracle_grid({"rec_xyz": xyz, "rec_types": types}, sfp, origin, n, np.arange(18))
    perm = np.random.default_rng(0).permutation(len(xyz))
    shuf = oracle_grid({"rec_xyz": xyz[perm], "rec_types": types[perm]}, sfp, origin, n, np.arange(18))
    np.testing.assert_allclose(shuf, full, rtol=1e-12, atol=1e-12)
    sub = oracle_grid({"rec_xyz": xyz, "rec_types": types}, sfp, origin, n, [0, 3, 7])
    np.testing.assert_array_equal(sub, full[[0, 3, 7]])
    pt = oracle_grid({"rec_xyz": xyz, "rec_types": types}, sfp, origin, n, [0, 3, 7], per_term=True)
    w = np.array([sfp[k] for k in ("w_g1", "w_g2", "w_rep", "w_hyd", "w_hb")], np.float32).astype(np.float64)
    np.testing.assert_allclose(np.einsum("m,mq...->q...", w, pt), sub, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------------------------------------ pose energy
def _small_problem(n_heavy=14, n_tors=3, seed=5, grid=None, box=16.0):
    rec = synth.make_receptor(seed, 400, 13.0, 5.0, min_sep=3.0)
    lig = synth.make_ligand(seed + 1, n_heavy, n_tors)
    prob, meta = cpu_problem(rec, [lig], np.zeros(3), box, grid=grid)
    return rec, lig, prob, meta


def _rot_inc(s, rv):
    """state with q <- quat(rv) * q (the pose-space rotation the gradient's torque component refers to)."""
    s = s.astype(np.float64).copy()
    ang = np.linalg.norm(rv)
    k = math.sin(ang / 2) / ang if ang > 0 else 0.5
    a = np.array([math.cos(ang / 2), *(rv * k)])
    b = s[3:7]
    q = np.array([a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3],
                  a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2],
                  a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1],
                  a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0]])
    s[3:7] = q / np.linalg.norm(q)
    return s


def test_gradient_vs_central_differences():
    """Analytic FP64 gradient (translation, torque, torsions) = central differences of the FP64 energy, on poses whose
    atoms and pairs all sit further than 0.05 cell from any discontinuity (cell faces, cutoff, slope kinks). States
    cross the interface as float32, so each difference quotient divides by the float32 step actually taken."""
    rec, lig, prob, meta = _small_problem()
    st = states(11, 96, np.zeros(3), 10.0, lig, prob.KM)
    e, g, X, m = prob.energy_grad(np.zeros(len(st), np.int32), st)
    h, hr = 1e-4, 2e-4          # A, rad (rotation moves atoms by <= 6 A * hr)
    checked = 0
    for i in range(len(st)):
        if m[i] < 2.5e-3:
            continue
        s0 = st[i]
        fd = np.zeros(6 + lig.n_tors)
        for d in list(range(3)) + [7 + k for k in range(lig.n_tors)]:
            sp, sm = s0.copy(), s0.copy()
            sp[d] += np.float32(h); sm[d] -= np.float32(h)
            fd[d if d < 3 else d - 1] = (_e(prob, sp) - _e(prob, sm)) / (float(sp[d]) - float(sm[d]))
        for d in range(3):
            rv = np.zeros(3); rv[d] = hr
            fd[3 + d] = (_e(prob, _rot_inc(s0, rv)) - _e(prob, _rot_inc(s0, -rv))) / (2 * hr)
        ga = g[i, :6 + lig.n_tors]
        assert np.linalg.norm(ga - fd) <= 2e-3 * max(np.linalg.norm(fd), 1.0), (i, ga, fd)
        checked += 1
    assert checked >= 20


def _e(prob, s):
    e, _, _, _ = prob.energy_grad(np.zeros(1, np.int32), np.asarray(s, np.float32)[None])
    return e[0]


def test_linear_field_is_exact():
    """With every grid filled with f(x) = a.x + b, trilinear lookup is exact, so with no intra pairs
    E = sum_a (a.X_a + b), dE/dt = n a, torque = sum_a (X_a - t) x a."""
    a, b = np.array([0.3, -0.2, 0.5]), 0.7

    def lin(origin, n, types):
        ii = np.stack(np.meshgrid(*[np.arange(k) for k in n], indexing="ij"), -1).astype(np.float64)
        pos = origin + ii * SPACING
        v = (pos @ a + b).astype(np.float32)
        return np.broadcast_to(v, (len(types),) + tuple(n)).copy()

    rec, lig, prob, meta = _small_problem(grid=lin, box=24.0)
    prob.lig["np"][:] = 0
    st = states(3, 32, np.zeros(3), 8.0, lig, prob.KM)
    e, g, X, _ = prob.energy_grad(np.zeros(len(st), np.int32), st)
    na = int(prob.lig["na"][0])
    for i in range(len(st)):
        Xi = X[i, :na]
        # grid values are float32-rounded samples of f, so the expected sum carries their rounding only
        np.testing.assert_allclose(e[i], (Xi @ a + b).sum(), rtol=0, atol=2e-5 * na)
        np.testing.assert_allclose(g[i, :3], na * a, rtol=0, atol=1e-4)
        np.testing.assert_allclose(g[i, 3:6], np.cross(Xi - st[i, :3].astype(np.float64), a).sum(0), atol=1e-4)


def test_intra_energy_matches_pairwise_closed_form_and_is_rigid_invariant():
    zero = lambda origin, n, types: np.zeros((len(types),) + tuple(n), np.float32)
    rec, lig, prob, meta = _small_problem(n_heavy=24, n_tors=5, grid=zero, box=40.0)
    st = states(4, 16, np.zeros(3), 20.0, lig, prob.KM)
    e, g, X, _ = prob.energy_grad(np.zeros(len(st), np.int32), st)
    typ = prob.lig["typ"][0]
    for i in range(len(st)):
        terms = [analytic_pair("vina", float(np.linalg.norm(X[i, p] - X[i, q])), int(typ[p]), int(typ[q]))
                 for p, q in prob.lig["pr"][0, :prob.lig["np"][0]] if p != q]
        want = sum(terms)
        assert abs(e[i] - want) <= 2e-7 * sum(abs(x) for x in terms) + 1e-9   # float32 weights
        # a different rigid placement of the same conformation has the same intra energy and zero net force
        s2 = st[i].copy()
        s2[:3] += np.float32(3.0)
        s2[3:7] = _rot_inc(st[i], np.array([0.4, -1.1, 0.3]))[3:7]
        e2, g2, _, _ = prob.energy_grad(np.zeros(1, np.int32), s2[None])
        assert abs(e2[0] - e[i]) <= 1e-8 * max(1, abs(e[i]))
        assert np.abs(g2[0, :3]).max() < 1e-9 and np.abs(g2[0, 3:6]).max() < 1e-8


def test_fk_preserves_bonds_and_sets_dihedrals():
    zero = lambda origin, n, types: np.zeros((len(types),) + tuple(n), np.float32)
    rec, lig, prob, meta = _small_problem(n_heavy=18, n_tors=4, grid=zero, box=40.0)
    s0 = np.zeros((1, 7 + prob.KM), np.float32)
    s0[0, 3] = 1.0
    _, _, X0, _ = prob.energy_grad(np.zeros(1, np.int32), s0)
    np.testing.assert_allclose(X0[0, :lig.n], lig.ref, atol=1e-6)
    nbr = [[] for _ in range(lig.n)]
    for i, j in lig.bonds:
        nbr[i].append(j); nbr[j].append(i)
    rng = np.random.default_rng(0)
    for trial in range(5):
        s = s0.copy()
        th = rng.uniform(-math.pi, math.pi, lig.n_tors).astype(np.float32)
        s[0, 7:7 + lig.n_tors] = th
        _, _, X, _ = prob.energy_grad(np.zeros(1, np.int32), s)
        X = X[0, :lig.n]
        for i, j in lig.bonds:
            assert abs(np.linalg.norm(X[i] - X[j]) - np.linalg.norm(lig.ref[i] - lig.ref[j])) < 1e-6
        for c in range(lig.n):   # 1-3 distances
            for i in nbr[c]:
                for j in nbr[c]:
                    if i < j:
                        assert abs(np.linalg.norm(X[i] - X[j]) - np.linalg.norm(lig.ref[i] - lig.ref[j])) < 1e-5
        for k, (a, b) in enumerate(lig.tors_axis):   # every dihedral across a rotatable bond moves by theta_k
            i = next(n for n in nbr[a] if n != b)
            l = next(n for n in nbr[b] if n != a)
            d0, d1 = _dihedral(lig.ref, i, a, b, l), _dihedral(X, i, a, b, l)
            diff = (d1 - d0 - float(th[k]) + math.pi) % (2 * math.pi) - math.pi
            assert abs(diff) < 1e-5, (k, d0, d1, th[k])


def _dihedral(P, i, j, k, l):
    b0, b1, b2 = P[i] - P[j], P[k] - P[j], P[l] - P[k]
    b1n = b1 / np.linalg.norm(b1)
    v = b0 - np.dot(b0, b1n) * b1n
    w = b2 - np.dot(b2, b1n) * b1n
    return math.atan2(np.dot(np.cross(b1n, v), w), np.dot(v, w))


def test_penalty_closed_form():
    zero = lambda origin, n, types: np.zeros((len(types),) + tuple(n), np.float32)
    rec, lig, prob, meta = _small_problem(grid=zero, box=8.0)
    prob.lig["np"][:] = 0
    s = np.zeros((1, 7 + prob.KM), np.float32)
    s[0, 3] = 1.0
    s[0, :3] = [11.0, -2.0, 0.5]
    e, g, X, _ = prob.energy_grad(np.zeros(1, np.int32), s)
    lo, hi = -4.0, 4.0
    Xa = X[0, :lig.n]
    over = np.maximum(lo - Xa, 0) ** 2 + np.maximum(Xa - hi, 0) ** 2
    assert e[0] == pytest.approx(PENALTY * over.sum(), rel=1e-12)
    grad = (2 * PENALTY * (np.maximum(Xa - hi, 0) - np.maximum(lo - Xa, 0))).sum(0)
    np.testing.assert_allclose(g[0, :3], grad, rtol=1e-12)


# ------------------------------------------------------------------------------------------------ BFGS / MC
def test_bfgs_first_iteration_on_quadratic_penalty():
    """One atom far outside the box on a zero grid: E = pen |x - hi|^2 is an exact quadratic in translation, so one
    BFGS iteration is a closed form: p = -g, Armijo halvings until E(x + a p) - E(x) < 1e-4 a p.g."""
    zero = lambda origin, n, types: np.zeros((len(types),) + tuple(n), np.float32)
    rec = synth.make_receptor(1, 50, 8.0, 3.0)
    lig = synth.make_ligand(2, 1, 0)
    prob, _ = cpu_problem(rec, [lig], np.zeros(3), 8.0, grid=zero)
    x0 = np.array([6.0, 0.0, 0.0])
    s = np.zeros((1, 7 + prob.KM), np.float32)
    s[0, :3], s[0, 3] = x0, 1.0
    cur, best = s.copy(), s.copy()
    ce, be = np.zeros(1, np.float32), np.full(1, np.inf, np.float32)
    prob.run_mc([0], [0], cur, ce, best, be, True, 0, 1, 1.2, 2.0, 2.0, 0, 0)
    E = lambda x: PENALTY * (max(x - 4.0, 0.0) ** 2 + max(-4.0 - x, 0.0) ** 2)
    g = 2 * PENALTY * (x0[0] - 4.0)
    a = 1.0
    while not (E(x0[0] - a * g) - E(x0[0]) < 1e-4 * a * (-g * g)):
        a *= 0.5
    want = E(x0[0] - a * g)
    assert ce[0] == pytest.approx(want, rel=1e-6, abs=1e-6)
    assert cur[0, 0] == pytest.approx(x0[0] - a * g, rel=1e-6, abs=1e-6)


def test_bfgs_converges_on_quadratic():
    zero = lambda origin, n, types: np.zeros((len(types),) + tuple(n), np.float32)
    rec = synth.make_receptor(1, 50, 8.0, 3.0)
    lig = synth.make_ligand(2, 1, 0)
    prob, _ = cpu_problem(rec, [lig], np.zeros(3), 8.0, grid=zero)
    s = np.zeros((1, 7 + prob.KM), np.float32)
    s[0, :3], s[0, 3] = [7.0, -9.0, 5.5], 1.0
    cur, best = s.copy(), s.copy()
    ce, be = np.zeros(1, np.float32), np.full(1, np.inf, np.float32)
    prob.run_mc([0], [0], cur, ce, best, be, True, 0, 30, 1.2, 2.0, 2.0, 0, 0)
    assert ce[0] < 1e-6 and be[0] < 1e-6


@pytest.fixture(scope="module")
def small_cx():
    w = synth.make_single("small", 0)
    rec = w["receptor"]
    prob, meta = cpu_problem(rec, [w["lig"]], w["center"], w["size"])
    return w, prob


def test_mc_monotone_and_best_re_evaluates(small_cx):
    w, prob = small_cx
    lig = w["lig"]
    N = 8
    st = states(0, N, w["center"], w["size"], lig, prob.KM)
    e0, _, _, _ = prob.energy_grad(np.zeros(N, np.int32), st)
    cur, best = st.copy(), st.copy()
    ce, be = np.zeros(N, np.float32), np.full(N, np.inf, np.float32)
    ev = prob.run_mc(np.zeros(N), np.arange(N), cur, ce, best, be, True, 5, 0, 1.2, 2.0, 2.0, 12345, 0)
    assert (be <= e0 + 1e-4).all()
    assert (be <= ce + 1e-6).all()
    e1, _, _, _ = prob.energy_grad(np.zeros(N, np.int32), best)
    np.testing.assert_allclose(e1, be, rtol=1e-4, atol=1e-4)
    assert (ev > 5).all()


def test_mc_segments_equal_one_launch(small_cx):
    """Counter-based RNG: 2 segments of 3 steps (step_offset 0, 3) == one launch of 6 steps, bit for bit, in the
    float32 restatement (its state never leaves float32, so the segment boundary is lossless)."""
    w, prob = small_cx
    N = 4
    st = states(1, N, w["center"], w["size"], w["lig"], prob.KM)
    outs = []
    for split in (False, True):
        cur, best = st.copy(), st.copy()
        ce, be = np.zeros(N, np.float32), np.full(N, np.inf, np.float32)
        if split:
            prob.run_mc(np.zeros(N), np.arange(N), cur, ce, best, be, True, 3, 0, 1.2, 2.0, 2.0, 99, 0, f64=False)
            prob.run_mc(np.zeros(N), np.arange(N), cur, ce, best, be, False, 3, 0, 1.2, 2.0, 2.0, 99, 3, f64=False)
        else:
            prob.run_mc(np.zeros(N), np.arange(N), cur, ce, best, be, True, 6, 0, 1.2, 2.0, 2.0, 99, 0, f64=False)
        outs.append((cur.copy(), ce.copy(), best.copy(), be.copy()))
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


def test_exact_mode_matches_direct_pairwise(small_cx):
    """exact=1: E = sum over (ligand atom, receptor atom) pairs of the two-atom energy + penalty + intra."""
    w, prob = small_cx
    N = 6
    st = states(2, N, w["center"], w["size"], w["lig"], prob.KM, sigma=1.0)
    e, g, X, _ = prob.energy_grad(np.zeros(N, np.int32), st, exact=True)
    zero_intra = dict(prob.lig)
    rec = w["receptor"]
    typ = prob.lig["typ"][0]
    na = int(prob.lig["na"][0])
    for i in range(N):
        terms = []
        for a in range(na):
            r = np.linalg.norm(rec["rec_xyz"].astype(np.float64) - X[i, a], axis=1)
            terms += [analytic_pair("vina", rr, int(typ[a]), int(tb)) for rr, tb in zip(r, rec["rec_types"]) if rr < 8]
        terms += [analytic_pair("vina", float(np.linalg.norm(X[i, p] - X[i, q])), int(typ[p]), int(typ[q]))
                  for p, q in prob.lig["pr"][0, :prob.lig["np"][0]] if p != q]
        lo, hi = -10.0, 10.0
        pen = PENALTY * (np.maximum(lo - X[i, :na], 0) ** 2 + np.maximum(X[i, :na] - hi, 0) ** 2).sum()
        assert abs(e[i] - (sum(terms) + pen)) <= 2e-7 * sum(abs(x) for x in terms) + 1e-9
