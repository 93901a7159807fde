"""TEST INFRASTRUCTURE ONLY: ctypes binding of the reference's own host kernels built by oracle/build_ref.py.

RefKernels(AM, KM) takes the same packed numpy arrays as oracle.OracleProblem (adock_amd.ligand.pack_ligands output and
the grid / scoring / receptor dicts) and calls the reference's kernels in their parameter order: ligand (13), grid
(10), scoring function (5), receptor (4), exact, then the kernel's own (src/adock/mckernel.py:328-345,411-425,451-458;
host launchers as src/adock/cpukernel.py:97-101,104-108). Arithmetic is the reference host build's: FP32 with
-ffast-math. Only tests/ and bench.py's CPU-baseline leg use it, as a checker or a timed baseline.
"""
import ctypes as C

import numpy as np

from . import build_ref

P, I, F, LL, U64 = C.c_void_p, C.c_int, C.c_float, C.c_longlong, C.c_uint64
_LIG = [P] * 11 + [I, P]                 # ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot
_GRID = [P] * 7 + [I, F, F]              # grid, origin, lo, hi, rec_idx, goffs, gdims, T, spacing, pen
_SF = [P] * 5
_REC = [P, P, P, I]


def _ptr(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], type(a)
    return a.ctypes.data


_LIBS = {}


def available(AM=64, KM=8):
    return build_ref.available(AM, KM)


def _load(AM, KM):
    key = (AM, KM)
    if key not in _LIBS:
        path = build_ref.lib_path(AM, KM)
        if not path.exists():
            build_ref.build(buckets=[key])
        L = C.CDLL(str(path))
        L.ref_mc_kernel.argtypes = ([LL] + _LIG + _GRID + _SF + _REC + [I]
                                    + [P, P, I, P, P, I, P, P, P, P, I, I, I, F, F, F, U64, I])
        L.ref_eg_kernel.argtypes = [LL] + _LIG + _GRID + _SF + _REC + [I] + [P, I, P, P, P]
        L.ref_stagea_kernel.argtypes = [LL, P, P, P, I, P, I, P, P, I] + [P] * 7 + [F, F, F, P, P]
        L.launch_build_grid_tiled.argtypes = [P, P, I] + [P] * 5 + [I] + [F] * 4 + [I] * 3 + [P, I, P, I]
        L.set_threads.argtypes = [I]
        L.get_threads.restype = I
        for f in (L.ref_mc_kernel, L.ref_eg_kernel, L.ref_stagea_kernel, L.launch_build_grid_tiled, L.set_threads):
            f.restype = None
        _LIBS[key] = L
    return _LIBS[key]


class RefKernels:
    def __init__(self, AM, KM, lig, grid, sf, rec=None):
        self.lib = _load(AM, KM)
        self.AM, self.KM, self.SD, self.DM = AM, KM, 7 + KM, 6 + KM
        c = np.ascontiguousarray
        self.lig = {k: (c(v) if isinstance(v, np.ndarray) else v) for k, v in lig.items()}
        self.grid = {k: (c(v) if isinstance(v, np.ndarray) else v) for k, v in grid.items()}
        self.sf = {k: c(v) for k, v in sf.items()}
        if rec is None:
            rec = {"rxyz": np.zeros((1, 1, 3), np.float32), "rtyp": np.zeros((1, 1), np.int32),
                   "rn": np.zeros(1, np.int32)}
        self.rec = {k: c(v) for k, v in rec.items()}

    def set_threads(self, n):
        self.lib.set_threads(int(n))

    def _common(self, exact):
        L, G, S, R = self.lig, self.grid, self.sf, self.rec
        return ([_ptr(L[k]) for k in ("ref", "typ", "tm", "ax", "axp", "axu", "pr", "na", "nt", "np", "rg")]
                + [int(L["P"]), _ptr(L["gslot"])]
                + [_ptr(G[k]) for k in ("flat", "origin", "lo", "hi", "rec_idx", "offs", "dims")]
                + [0, float(G["spacing"]), float(G["penalty"])]
                + [_ptr(S[k]) for k in ("prm", "rad", "hyd", "don", "acc")]
                + [_ptr(R["rxyz"]), _ptr(R["rtyp"]), _ptr(R["rn"]), int(R["rxyz"].shape[1]), int(exact)])

    def energy_grad(self, chain_lig, state, exact=False):
        """-> (e [N] f32, g [N, DM] f32) from the reference's eg_kernel."""
        cl = np.ascontiguousarray(chain_lig, np.int32)
        st = np.ascontiguousarray(state, np.float32)
        n = len(cl)
        e = np.zeros(n, np.float32)
        g = np.zeros((n, self.DM), np.float32)
        self.lib.ref_eg_kernel(n, *self._common(exact), _ptr(cl), n, _ptr(st), _ptr(e), _ptr(g))
        return e, g

    def run_mc(self, chain_lig, chain_id, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos,
               amp_rot, seed, step_offset, exact=False, nbr=None, nbr_n=None):
        """In place on float32 numpy cur/cur_e/best/best_e (the reference's mc_kernel, one OpenMP task per chain)."""
        for a in (cur, cur_e, best, best_e):
            assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
        cl = np.ascontiguousarray(chain_lig, np.int32)
        ci = np.ascontiguousarray(chain_id, np.int32)
        nb = np.ascontiguousarray(nbr, np.int32) if nbr is not None else None
        nbn = np.ascontiguousarray(nbr_n, np.int32) if nbr_n is not None else None
        n = len(cl)
        self.lib.ref_mc_kernel(n, *self._common(exact), _ptr(cl), _ptr(ci), n, _ptr(nb), _ptr(nbn),
                               int(nb.shape[1]) if nb is not None else 0, _ptr(cur), _ptr(cur_e), _ptr(best),
                               _ptr(best_e), int(init_opt), int(mc_steps), int(bfgs_iters), float(temp),
                               float(amp_pos), float(amp_rot), int(seed) & (2 ** 64 - 1), int(step_offset))

    def stage_a(self, conf, quats, trans, ntrans, cap):
        """-> (best_e [L, C*Q] f32, best_i [L, C*Q] int32) from the reference's stagea_kernel."""
        conf = np.ascontiguousarray(conf, np.float32)
        quats = np.ascontiguousarray(quats, np.float32)
        trans = np.ascontiguousarray(trans, np.float32)
        ntrans = np.ascontiguousarray(ntrans, np.int32)
        L, Cn = conf.shape[:2]
        Q, NT = quats.shape[0], trans.shape[1]
        be = np.zeros(L * Cn * Q, np.float32)
        bi = np.zeros(L * Cn * Q, np.int32)
        G = self.grid
        self.lib.ref_stagea_kernel(L * Cn * Q, _ptr(conf), _ptr(self.lig["gslot"]), _ptr(self.lig["na"]), Cn,
                                   _ptr(quats), Q, _ptr(trans), _ptr(ntrans), NT,
                                   *[_ptr(G[k]) for k in ("flat", "origin", "lo", "hi", "rec_idx", "offs", "dims")],
                                   float(G["spacing"]), float(G["penalty"]), float(cap), _ptr(be), _ptr(bi))
        return be.reshape(L, Cn * Q), bi.reshape(L, Cn * Q)


def build_grid_tiled(rec_xyz, rec_types, sfp_arrays, prm_grid, origin, spacing, n, lig_types, tile=8, threads=None):
    """The reference's host grid builder (launch_build_grid_tiled, src/adock/cpukernel.py:104-137) -> [T, nx, ny, nz]
    float32. sfp_arrays = (radius f32, hyd, don, acc int32); prm_grid: 15 floats in grid order."""
    L = _load(64, 8)
    if threads:
        L.set_threads(int(threads))
    nx, ny, nz = map(int, n)
    rx = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
    rt = np.ascontiguousarray(rec_types, np.int32)
    lt = np.ascontiguousarray(lig_types, np.int32)
    rad, hyd, don, acc = [np.ascontiguousarray(a, t) for a, t in zip(sfp_arrays, (np.float32, np.int32, np.int32,
                                                                                  np.int32))]
    prm = np.ascontiguousarray(prm_grid, np.float32)
    out = np.zeros(len(lt) * nx * ny * nz, np.float32)
    L.launch_build_grid_tiled(_ptr(rx), _ptr(rt), len(rx), _ptr(rad), _ptr(hyd), _ptr(don), _ptr(acc), _ptr(lt),
                              len(lt), float(origin[0]), float(origin[1]), float(origin[2]), float(spacing), nx, ny,
                              nz, _ptr(prm), 0, _ptr(out), int(tile))
    return out.reshape(len(lt), nx, ny, nz)
