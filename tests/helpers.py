"""Shared test helpers: independent closed forms, small synthetic problems, CPU oracle problems."""
import math

import numpy as np

import oracle
from adock_amd import synth
from adock_amd.config import grid_prm, load_config, search_prm, sf_params
from adock_amd.ligand import pack_ligands

CFG = load_config()
IDX = CFG["_xs_index"]
SPACING = CFG["grid"]["spacing"]
PENALTY = CFG["grid"]["out_of_box_penalty"]

# tolerances of the parity contract
GRID_REL = 1e-5          # grids vs the FP64 oracle, relative with an absolute floor
GRID_FLOOR = 1e-5        # kcal/mol
EG_REL = 1e-4            # per-pose energies and gradients vs the FP64 oracle
EG_FLOOR = 1e-4


def analytic_pair(sf, r, a, b):
    """Two-atom energy from the published Vina / Vinardo form (written independently of the kernels and oracle)."""
    p = CFG["scoring"][sf]
    tab = dict(zip(CFG["xs_types"]["_columns"], zip(*[CFG["xs_types"][n] for n in CFG["_xs_names"]])))
    R = lambda t: np.float32(tab[p["radius_column"]][t])
    if r >= p["cutoff"]:
        return 0.0
    d = r - float(R(a)) - float(R(b))
    e = p["gauss1"]["weight"] * math.exp(-((d - p["gauss1"]["offset"]) / p["gauss1"]["width"]) ** 2)
    e += p["gauss2"]["weight"] * math.exp(-((d - p["gauss2"]["offset"]) / p["gauss2"]["width"]) ** 2)
    e += p["repulsion"]["weight"] * (d * d if d < 0 else 0.0)
    if tab["hydrophobic"][a] and tab["hydrophobic"][b]:
        g, bd = p["hydrophobic"]["good"], p["hydrophobic"]["bad"]
        e += p["hydrophobic"]["weight"] * (1.0 if d < g else 0.0 if d > bd else (bd - d) / (bd - g))
    if (tab["donor"][a] and tab["acceptor"][b]) or (tab["acceptor"][a] and tab["donor"][b]):
        g, bd = p["hbond"]["good"], p["hbond"]["bad"]
        e += p["hbond"]["weight"] * (1.0 if d < g else 0.0 if d > bd else (bd - d) / (bd - g))
    return e


def sf_tables(sfp):
    return (np.asarray(sfp["radius"], np.float32), np.asarray(sfp["hydrophobic"], np.int32),
            np.asarray(sfp["donor"], np.int32), np.asarray(sfp["acceptor"], np.int32))


def random_receptor(n=300, seed=0, extent=14.0):
    """The reference's known-answer receptor recipe (tests/test_scoring_grid.py:54-58)."""
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(0, extent, (n, 3)).astype(np.float32)
    types = rng.choice([IDX[t] for t in ("C_H", "C_P", "N_D", "N_A", "O_A", "O_DA", "S_P", "Met_D")], n)
    return xyz, types.astype(np.int32)


def oracle_grid(rec, sfp, origin, n, lig_types, f64=True, per_term=False):
    return oracle.build_grid(rec["rec_xyz"], rec["rec_types"], sf_tables(sfp), grid_prm(sfp), origin, SPACING, n,
                             lig_types, per_term=per_term, f64=f64)


def cpu_problem(rec, preps, center, size, sf="vina", grid=None, rec_idx=None):
    """OracleProblem over one receptor with grids built by the FP64 oracle (rounded to float32, as the kernels store
    them) for the ligands' types + C_H. grid: optional callable(origin, n, types) -> [T, nx, ny, nz] array."""
    sfp = sf_params(CFG, sf)
    types = np.unique(np.concatenate([[0]] + [p.types[p.types >= 0] for p in preps]))
    slot = -np.ones((1, len(sfp["radius"])), np.int64)
    slot[0, types] = np.arange(len(types))
    center = np.asarray(center, float)
    size = np.broadcast_to(np.asarray(size, float), (3,))
    origin = center - size / 2
    n = tuple(int(math.ceil(s / SPACING)) + 1 for s in size)
    G = grid(origin, n, types) if grid is not None else oracle_grid(rec, sfp, origin, n, types)
    G = np.asarray(G, np.float32)
    rec_idx = np.zeros(len(preps), np.int64) if rec_idx is None else rec_idx
    AM, KM, lig = pack_ligands(preps, slot, rec_idx)
    gd = dict(flat=G.reshape(-1), origin=origin.astype(np.float32)[None], lo=origin.astype(np.float32)[None],
              hi=(center + size / 2).astype(np.float32)[None], rec_idx=rec_idx.astype(np.int32),
              offs=np.zeros(1, np.int64), dims=np.array([n], np.int64), spacing=SPACING, penalty=PENALTY)
    sft = dict(prm=search_prm(sfp), rad=sf_tables(sfp)[0], hyd=sf_tables(sfp)[1], don=sf_tables(sfp)[2],
               acc=sf_tables(sfp)[3])
    rd = dict(rxyz=np.asarray(rec["rec_xyz"], np.float32)[None], rtyp=np.asarray(rec["rec_types"], np.int32)[None],
              rn=np.array([len(rec["rec_xyz"])], np.int32))
    return oracle.OracleProblem(AM, KM, lig, gd, sft, rd), dict(G=G, origin=origin, n=n, types=types, slot=slot)


def states(seed, n, center, size, prep, KM, sigma=None):
    """Random poses: translation uniform in the shrunk box (or N(center, sigma)), uniform quaternion, theta U(-pi, pi)."""
    t, q, th = synth.start_states(seed, n, center, size, prep.n_tors, KM)
    if sigma is not None:
        rng = np.random.default_rng(seed + 7)
        t = (np.asarray(center) + rng.standard_normal((n, 3)) * sigma).astype(np.float32)
    s = np.zeros((n, 7 + KM), np.float32)
    s[:, :3], s[:, 3:7], s[:, 7:] = t, q, th
    return s


def rel_err(a, b, floor, rel):
    """max |a - b| / max(|b|, floor / rel) (the reference's _rel_close, tests/test_cpukernel.py:210-213)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float((np.abs(a - b) / np.maximum(np.abs(b), floor / rel)).max()) if a.size else 0.0


def vec_rel(a, b, floor):
    """max over rows of ||a - b|| / max(||b||, floor) (tests/test_cpukernel.py:203-207)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float((np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), floor)).max()) if a.size else 0.0


class OracleKernel:
    """CPU twin of adock_amd.kernel.KernelProblem for tests: same constructor and methods, every computation done by
    the FP64 oracle on host tensors (the role the reference's CpuKernelProblem plays, src/adock/cpukernel.py:231-306).
    Lets the search driver run on CPU for control-flow checks and as the comparison run for the GPU. The MC chains run
    in the FP32 restatement, which (like the kernels) keeps the state in float32 between launches, so a run split
    into segments equals the unsplit run."""

    def __init__(self, prob):
        import torch
        from adock_amd import workload
        self.AM, self.KM, lig, grid, sf, rec = workload.host_arrays(prob)
        self.op = oracle.OracleProblem(self.AM, self.KM, lig, grid, sf, rec)
        self.L, self.SD, self.DM = prob.lb.L, 7 + self.KM, 6 + self.KM
        self.device = torch.device("cpu")

    def pack(self, t, q, theta):
        from adock_amd.kernel import KernelProblem
        return KernelProblem.pack(self, t.cpu(), q.cpu(), theta.cpu())

    def unpack(self, s, K):
        from adock_amd.kernel import KernelProblem
        return KernelProblem.unpack(self, s, K)

    def energy_grad(self, chain_lig, state, exact=False):
        import torch
        e, g, _, _ = self.op.energy_grad(chain_lig.cpu().numpy(), state.cpu().numpy(), exact=exact)
        return torch.as_tensor(e, dtype=torch.float32), torch.as_tensor(g, dtype=torch.float32)

    def stage_a(self, conf_h, quats, trans, ntrans, cap):
        import torch
        be, bi = self.op.stage_a(conf_h.cpu().numpy(), quats.cpu().numpy(), trans.cpu().numpy(),
                                 ntrans.cpu().numpy(), cap)
        return torch.as_tensor(be, dtype=torch.float32), torch.as_tensor(bi, dtype=torch.int64)

    def receptor_neighbours(self, prob, chain_lig, states, K, margin):
        from adock_amd.kernel import KernelProblem
        return KernelProblem.receptor_neighbours(self, prob, chain_lig, states, K, margin)

    def run_mc(self, chain_lig, chain_id, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos,
               amp_rot, seed, step_offset, exact=False, nbr=None, nbr_n=None):
        self.op.run_mc(chain_lig.cpu().numpy(), chain_id.cpu().numpy(), cur.numpy(), cur_e.numpy(), best.numpy(),
                       best_e.numpy(), init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot, seed, step_offset,
                       exact=exact, nbr=None if nbr is None else nbr.cpu().numpy(),
                       nbr_n=None if nbr_n is None else nbr_n.cpu().numpy(), f64=False)


def cpu_search_problem(rec, preps, center, size, sf="vina", rec_idx=None):
    """A search Problem on host tensors: oracle-built grids in a GridSet, the ligand batch, the pair potential and the
    padded receptor (for the exact polish)."""
    import torch
    from adock_amd.grid import GridSet
    from adock_amd.problem import LigandBatch, PairPotential, Problem, pad_receptors
    _, meta = cpu_problem(rec, preps, center, size, sf=sf, rec_idx=rec_idx)
    size = np.broadcast_to(np.asarray(size, float), (3,))
    center = np.asarray(center, float)
    G = torch.as_tensor(meta["G"])
    gs = GridSet([G], torch.as_tensor(meta["origin"], dtype=torch.float32)[None],
                 torch.as_tensor(center - size / 2, dtype=torch.float32)[None],
                 torch.as_tensor(center + size / 2, dtype=torch.float32)[None], SPACING, PENALTY,
                 slot=torch.as_tensor(meta["slot"]), n_types=G.shape[0])
    ri = torch.zeros(len(preps), dtype=torch.long) if rec_idx is None else torch.as_tensor(rec_idx)
    return Problem(gs, LigandBatch(preps, "cpu"), ri, PairPotential(sf_params(CFG, sf), "cpu"),
                   rec=pad_receptors([rec], "cpu"))


def fit_unit_weights(m):
    """Weights with 1 on term m (name or index) and 0 elsewhere."""
    from adock_amd.fit.terms import TERM_NAMES
    w = np.zeros(len(TERM_NAMES), np.float32)
    w[TERM_NAMES.index(m) if isinstance(m, str) else m] = 1.0
    return w


def cpu_fitted_problem(cxs, spec, w):
    """A fitted-function search Problem on host tensors: compact ligand types, FittedPotential, grids from the FP64
    oracle (one common cube per complex, the make_gridset_fit rule) and the padded receptors (pair types TL + ext)."""
    import torch
    from adock_amd.fit.fdock import compact_ligands
    from adock_amd.fit.terms import FittedPotential
    from adock_amd.grid import GridSet
    from adock_amd.problem import LigandBatch, Problem, pad_receptors
    lig_list, ligs = compact_ligands(cxs)
    pot = FittedPotential(spec, w, lig_list, "cpu")
    n = int(math.ceil(max(float(np.max(c["size"])) for c in cxs) / SPACING)) + 1
    grids, org, lo, hi = [], [], [], []
    for c in cxs:
        o = np.asarray(c["center"], float) - np.asarray(c["size"], float) / 2
        grids.append(oracle.build_grid_fit(c["rec_xyz"], c["rec_ext"], spec.rec_radius, pot.radius.numpy(),
                                           pot.W.numpy(), spec.grid_prm(), o, SPACING, n).astype(np.float32))
        org.append(o)
        lo.append(o)
        hi.append(np.asarray(c["center"], float) + np.asarray(c["size"], float) / 2)
    T = lambda a: torch.as_tensor(np.array(a), dtype=torch.float32)
    gs = GridSet(torch.as_tensor(np.stack(grids)), T(org), T(lo), T(hi), SPACING, PENALTY)
    recs = [{"rec_xyz": c["rec_xyz"], "rec_types": np.asarray(c["rec_ext"], np.int64) + pot.TL} for c in cxs]
    return Problem(gs, LigandBatch(ligs, "cpu"), torch.arange(len(cxs)), pot, rec=pad_receptors(recs, "cpu"))
