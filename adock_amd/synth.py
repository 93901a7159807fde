"""Seeded synthetic docking workloads: receptors, ligands with torsion trees, boxes and chain start states.

Receptor: heavy atoms on a jittered, randomly rotated FCC lattice inside a sphere, with an empty spherical cavity at the
box centre. Random sequential placement saturates near 0.027 atoms/A^3 at 3.0 A separation, below the 0.046-0.056
atoms/A^3 the workloads ask for, so a lattice whose nearest-neighbour spacing exceeds the minimum separation by twice
the jitter amplitude is used instead (every pair stays at least `min_sep` apart by construction).
XS type mix: C 63% (half C_H, half C_P), N 17%, O 19%, S 1%; N and O get donor / acceptor flags drawn per atom
(N: donor 0.5, acceptor 0.3; O: donor 0.3, acceptor 0.8).

Ligand: heavy-atom random tree, 1.5 A bonds, 109.47 deg bond angle to a random existing bond of the parent, at most 3
bonds per atom, growth steps rejected when a new atom comes closer than 2.5 A to any atom that is not its parent or a
neighbour of its parent. Elements C 70 / N 12 / O 15 / S + halogen 3 (halogens only on terminal atoms, O and S at most
2 bonds). Rotatable bonds are drawn among non-terminal bonds (capped at their number).
"""
import math

import numpy as np

from .config import load_config
from .ligand import LigandPrep, random_quat

_CFG = load_config()
IDX = _CFG["_xs_index"]


def _fcc_sites(a_nn, radius):
    a = a_nn * math.sqrt(2.0)
    m = int(math.ceil(radius / a)) + 2
    g = np.arange(-m, m + 1)
    base = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    offs = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    return ((base[:, None, :] + offs[None]) * a).reshape(-1, 3)


def _random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def receptor_types(rng, n):
    u = rng.random(n)
    t = np.empty(n, np.int32)
    for i in range(n):
        if u[i] < 0.315:
            t[i] = IDX["C_H"]
        elif u[i] < 0.63:
            t[i] = IDX["C_P"]
        elif u[i] < 0.80:
            d, a = rng.random() < 0.5, rng.random() < 0.3
            t[i] = IDX["N_DA" if d and a else "N_D" if d else "N_A" if a else "N_P"]
        elif u[i] < 0.99:
            d, a = rng.random() < 0.3, rng.random() < 0.8
            t[i] = IDX["O_DA" if d and a else "O_D" if d else "O_A" if a else "O_P"]
        else:
            t[i] = IDX["S_P"]
    return t


def make_receptor(seed, n_atoms, radius, cavity, min_sep=3.0):
    """Returns dict(rec_xyz [n,3] f32, rec_types [n] i32) centred on the origin."""
    rng = np.random.default_rng(seed)
    shell = 4.0 / 3.0 * math.pi * (radius ** 3 - cavity ** 3)
    # FCC number density is sqrt(2)/a_nn^3: pick the largest spacing that still leaves 3% spare sites
    a_nn = min(min_sep + 0.1, (math.sqrt(2.0) * shell / (1.03 * n_atoms)) ** (1.0 / 3.0))
    if a_nn <= min_sep:
        raise ValueError(f"{n_atoms} atoms at >= {min_sep} A do not fit a {radius} A sphere with a {cavity} A cavity")
    jitter = 0.49 * (a_nn - min_sep)
    sites = _fcc_sites(a_nn, radius) @ _random_rotation(rng).T + rng.uniform(-a_nn, a_nn, 3)
    r = np.linalg.norm(sites, axis=1)
    sites = sites[(r <= radius - jitter) & (r >= cavity + jitter)]
    if len(sites) < n_atoms:
        raise ValueError("lattice too small for the requested atom count")
    sites = sites[rng.choice(len(sites), n_atoms, replace=False)]
    v = rng.standard_normal((n_atoms, 3))
    v *= (jitter * rng.random(n_atoms) ** (1 / 3) / np.linalg.norm(v, axis=1))[:, None]
    xyz = (sites + v).astype(np.float32)
    return {"rec_xyz": xyz, "rec_types": receptor_types(rng, n_atoms)}


_MAXDEG = {"C": 3, "N": 3, "O": 2, "S": 2, "F": 1, "Cl": 1, "Br": 1}


def _grow_tree(rng, n):
    pos = np.zeros((n, 3))
    nbr = [[] for _ in range(n)]
    bonds = []
    cos_t = math.cos(math.radians(109.47))
    sin_t = math.sin(math.radians(109.47))
    k = 1
    fails = 0
    while k < n:
        cand = [i for i in range(k) if len(nbr[i]) < 3]
        p = int(rng.choice(cand))
        placed = False
        for _ in range(60):
            if not nbr[p]:
                d = rng.standard_normal(3)
                d /= np.linalg.norm(d)
            else:
                b = pos[nbr[p][rng.integers(len(nbr[p]))]] - pos[p]
                b /= np.linalg.norm(b)
                e1 = np.cross(b, [1.0, 0, 0] if abs(b[0]) < 0.9 else [0, 1.0, 0])
                e1 /= np.linalg.norm(e1)
                e2 = np.cross(b, e1)
                phi = rng.uniform(0, 2 * math.pi)
                d = cos_t * b + sin_t * (math.cos(phi) * e1 + math.sin(phi) * e2)
                ok = True
                for q in nbr[p]:
                    c = pos[q] - pos[p]
                    if np.dot(c, d) / np.linalg.norm(c) > math.cos(math.radians(100.0)):
                        ok = False
                        break
                if not ok:
                    continue
            x = pos[p] + 1.5 * d
            excl = set([p] + nbr[p])
            dist = np.linalg.norm(pos[:k] - x, axis=1)
            if all(dist[i] >= 2.5 for i in range(k) if i not in excl):
                pos[k] = x
                nbr[p].append(k)
                nbr[k].append(p)
                bonds.append((p, k))
                k += 1
                placed = True
                break
        if not placed:
            fails += 1
            if fails > 200:
                return None
    return pos, nbr, bonds


def _ligand_types(rng, nbr):
    n = len(nbr)
    elem = []
    for i in range(n):
        deg = len(nbr[i])
        for _ in range(20):
            u = rng.random()
            e = "C" if u < 0.70 else "N" if u < 0.82 else "O" if u < 0.97 else ("S" if rng.random() < 0.5 else
                                                                                 rng.choice(["F", "Cl", "Br"]))
            if deg <= _MAXDEG[e]:
                break
        else:
            e = "C"
        elem.append(str(e))
    types = np.empty(n, np.int64)
    for i in range(n):
        e, deg = elem[i], len(nbr[i])
        if e == "C":
            hetero = any(elem[j] in ("N", "O") for j in nbr[i])
            types[i] = IDX["C_P" if hetero else "C_H"]
        elif e == "N":
            u = rng.random()
            if deg == 1:
                types[i] = IDX["N_D" if u < 0.6 else "N_DA" if u < 0.8 else "N_A"]
            elif deg == 2:
                types[i] = IDX["N_A" if u < 0.5 else "N_D" if u < 0.8 else "N_P"]
            else:
                types[i] = IDX["N_P" if u < 0.7 else "N_D"]
        elif e == "O":
            types[i] = IDX["O_A" if deg == 2 or rng.random() < 0.6 else "O_DA"]
        elif e == "S":
            types[i] = IDX["S_P"]
        else:
            types[i] = IDX[{"F": "F_H", "Cl": "Cl_H", "Br": "Br_H"}[e]]
    return types


def make_ligand(seed, n_heavy, n_tors):
    """Random tree ligand as a LigandPrep (heavy atoms only)."""
    rng = np.random.default_rng(seed)
    for _ in range(100):
        g = _grow_tree(rng, n_heavy)
        if g is not None:
            break
    else:
        raise RuntimeError("ligand growth failed")
    pos, nbr, bonds = g
    types = _ligand_types(rng, nbr)
    inner = [(min(i, j), max(i, j)) for i, j in bonds if len(nbr[i]) > 1 and len(nbr[j]) > 1]
    k = min(n_tors, len(inner))
    rot = [inner[i] for i in sorted(rng.choice(len(inner), k, replace=False))] if k else []
    return LigandPrep(pos, types, bonds, rot)


def grid_dims(size, spacing=0.375):
    """Points per axis, the reference rule n = ceil(s / h) + 1 (src/adock/grid.py:272)."""
    return tuple(int(math.ceil(s / spacing)) + 1 for s in np.broadcast_to(np.asarray(size, float), (3,)))


def start_states(seed, n, center, size, n_tors, K):
    """Uniform translation in the box shrunk by 1 A, uniform quaternion, theta ~ U(-pi, pi) on real torsions.
    Returns (t [n,3], q [n,4], theta [n,K]) float32."""
    rng = np.random.default_rng(seed)
    c = np.asarray(center, np.float64)
    h = np.asarray(size, np.float64) / 2 - 1.0
    t = (c + rng.uniform(-h, h, (n, 3))).astype(np.float32)
    q = random_quat(rng, n)
    th = np.zeros((n, K), np.float32)
    th[:, :n_tors] = rng.uniform(-math.pi, math.pi, (n, n_tors))
    return t, q, th


# Workload definitions. Receptor radius / cavity / atom counts, ligand sizes, box edges, chains and steps.
WORKLOADS = {
    "small": dict(rec=dict(n_atoms=1500, radius=20.0, cavity=6.0, min_sep=3.0), lig=(20, 4), box=20.0, chains=8, steps=200,
            bfgs_iters=0),
    "single": dict(rec=dict(n_atoms=4000, radius=26.0, cavity=8.0, min_sep=2.8), lig=(35, 8), box=26.0, chains=256,
            steps=1000, bfgs_iters=0),
}


def make_single(name, seed=0):
    """A single-complex workload ("small": 1,500-atom receptor, 20-atom / 4-torsion ligand, 20 A box, 8 chains x 200 steps;
    "single": 4,000-atom receptor, 35-atom / 8-torsion ligand, 26 A box, 256 chains x 1,000 steps): receptor, ligand
    prep, box centre and size."""
    c = WORKLOADS[name]
    rec = make_receptor(seed, **c["rec"])
    lig = make_ligand(seed + 1, *c["lig"])
    return {"receptor": rec, "lig": lig, "center": np.zeros(3), "size": np.full(3, c["box"]), "chains": c["chains"],
            "steps": c["steps"], "bfgs_iters": c["bfgs_iters"]}


def make_screen(seed=0, n_ligands=512, chains=64, steps=500):
    """Screening batch: the GPU-size single-complex receptor and box (4,000 atoms around an 8 A cavity, 26 A box), and
    n_ligands ligands with heavy atoms ~ U[10, 60] and torsions ~ U[0, 15] (capped by the ligand's non-terminal
    bonds)."""
    rng = np.random.default_rng(seed + 991)
    rec = make_receptor(seed, **WORKLOADS["single"]["rec"])
    ligs = [make_ligand(seed * 100003 + 17 + i, int(rng.integers(10, 61)), int(rng.integers(0, 16)))
            for i in range(n_ligands)]
    return {"receptor": rec, "ligs": ligs, "center": np.zeros(3), "size": np.full(3, WORKLOADS["single"]["box"]),
            "chains": chains, "steps": steps, "bfgs_iters": 0}


def make_multi(seed=0, n_receptors=64, ligands_per=32, chains=32, steps=500):
    """Multi-receptor batch: n_receptors receptors of 2,000-6,000 atoms (sphere radius for ~0.05 atoms/A^3 around a 6-8 A
    cavity), box edges ~ U[16, 30] A, ligands_per ligands each from the U[10, 60] atoms / U[0, 15] torsions mix."""
    rng = np.random.default_rng(seed + 4242)
    recs, boxes, ligs, rec_idx = [], [], [], []
    for r in range(n_receptors):
        n = int(rng.integers(2000, 6001))
        cav = float(rng.uniform(6.0, 8.0))
        rad = (3.0 * n / (4.0 * math.pi * 0.05) + cav ** 3) ** (1.0 / 3.0)
        recs.append(make_receptor(seed * 7919 + r, n, rad, cav, min_sep=2.8))
        boxes.append(float(rng.uniform(16.0, 30.0)))
        for k in range(ligands_per):
            ligs.append(make_ligand(seed * 100003 + 50000 + r * ligands_per + k, int(rng.integers(10, 61)),
                                    int(rng.integers(0, 16))))
            rec_idx.append(r)
    return {"receptors": recs, "boxes": boxes, "ligs": ligs, "rec_idx": np.asarray(rec_idx), "chains": chains,
            "steps": steps, "bfgs_iters": 0}


def fitted_complex(rec, lig, center, size, spec, seed=0, p_charge=0.3, p_side_chain=0.5, rec_seed=None):
    """A complex for the fitted scoring function from a synthetic receptor and ligand: receptor buriedness from its own
    heavy atoms (count within spec.buried_radius), charge classes drawn on N (cation) and O (anion) atoms with
    probability p_charge (the annotation the reference derives from residue names and formal charges,
    src/adock/fit/terms.py:165-230), side-chain flags with probability p_side_chain. Returns the dict that
    adock_amd.fit.fdock takes (rec_xyz, rec_ext, center, size, lig, lig_ext). rec_seed: draw the receptor's
    annotation from its own seed, so that complexes of one pocket share one annotated receptor (a screen); None:
    from the complex's seed, like its ligand's classes."""
    from .fit.terms import buried_count, ligand_ext_types, receptor_ext_types
    rng = np.random.default_rng(seed + 77)
    rrng = rng if rec_seed is None else np.random.default_rng(rec_seed + 1077)
    names = _CFG["_xs_names"]
    elem = lambda t: names[t].split("_")[0]

    def classes(types, g):
        c = np.zeros(len(types), np.int64)
        for i, t in enumerate(types):
            if t >= 0 and g.random() < p_charge:
                c[i] = {"N": 1, "O": 2}.get(elem(int(t)), 0)
        return c

    xyz = np.asarray(rec["rec_xyz"], np.float32)
    ann = {"rec_types": np.asarray(rec["rec_types"], np.int64),
           "buried_count": buried_count(xyz, xyz, spec.buried_radius),
           "rec_chg": classes(rec["rec_types"], rrng), "rec_sc": rrng.random(len(xyz)) < p_side_chain}
    return {"rec_xyz": xyz, "rec_ext": receptor_ext_types(ann, spec), "center": np.asarray(center, np.float64),
            "size": np.broadcast_to(np.asarray(size, np.float64), (3,)).copy(), "lig": lig,
            "lig_ext": ligand_ext_types(lig.types, classes(lig.types, rng), spec.nxs)}


def make_native_complex(seed, n_heavy=20, n_tors=4, shell=(3.6, 4.6), outer=(4.6, 7.5), min_sep=3.0):
    """A complex with a known native pose: the ligand in its reference conformation at the origin (translation 0,
    identity rotation, torsions 0) and receptor atoms grown around it. Inner shell: points shell[0]..shell[1] A from
    the nearest ligand heavy atom, typed to complement it (O_A facing a ligand donor, N_D facing an acceptor, C_H
    otherwise); outer layer: C_H / C_P; every receptor pair at least min_sep apart. Returns dict(receptor, lig,
    center, native [n, 3] world coordinates)."""
    rng = np.random.default_rng(seed)
    lig = make_ligand(seed + 3, n_heavy, n_tors)
    X = lig.ref[lig.heavy].astype(np.float64)
    names = _CFG["_xs_names"]
    don = np.asarray(_CFG["_xs_table"]["donor"], bool)
    acc = np.asarray(_CFG["_xs_table"]["acceptor"], bool)
    lo, hi = X.min(0) - outer[1] - 1.0, X.max(0) + outer[1] + 1.0
    pts = rng.uniform(lo, hi, (40000, 3))
    d = np.sqrt(((pts[:, None, :] - X[None]) ** 2).sum(-1))
    dmin, near = d.min(1), d.argmin(1)
    keep = (dmin >= shell[0]) & (dmin <= outer[1])
    pts, dmin, near = pts[keep], dmin[keep], near[keep]
    order = np.argsort(dmin, kind="stable")            # inner points first
    chosen = []
    for i in order:
        p = pts[i]
        if chosen and np.min(np.sum((np.asarray(chosen) - p) ** 2, 1)) < min_sep ** 2:
            continue
        chosen.append(p)
    chosen = np.asarray(chosen)
    dc = np.sqrt(((chosen[:, None, :] - X[None]) ** 2).sum(-1))
    types = np.empty(len(chosen), np.int32)
    lt = lig.types[lig.heavy]
    for k in range(len(chosen)):
        j = int(dc[k].argmin())
        if dc[k, j] <= shell[1]:
            t = int(lt[j])
            types[k] = IDX["O_A"] if don[t] else IDX["N_D"] if acc[t] else IDX["C_H"]
        else:
            types[k] = IDX["C_H"] if rng.random() < 0.5 else IDX["C_P"]
    native = lig.ref.astype(np.float64)
    return {"receptor": {"rec_xyz": chosen.astype(np.float32), "rec_types": types}, "lig": lig,
            "center": np.zeros(3), "native": native, "xs_names": names}
