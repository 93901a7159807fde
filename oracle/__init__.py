"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (oracle.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this package, and only as the checker or
the timed CPU baseline; the product library never does.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_c = ctypes
P = _c.c_void_p
I = _c.c_int
F = _c.c_float
LL = _c.c_longlong
U64 = _c.c_ulonglong

_LIG = [P] * 11 + [I, P, I, I]          # ref..rg, P_MAX, gslot, AM, KM
_GRID = [P] * 7 + [F, F]               # grid..gdims, spacing, pen
_SF = [P] * 5
_REC = [P, P, P, I]


def build(force=False):
    """Compile liboracle.so / libcpubase.so with the committed Makefile."""
    if force or not (HERE / "liboracle.so").exists() or not (HERE / "libcpubase.so").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "-B" if force else "all"], check=True)


def _ptr(a):
    if a is None:
        return None
    assert isinstance(a, np.ndarray) and a.flags["C_CONTIGUOUS"], type(a)
    return a.ctypes.data


class _Lib:
    def __init__(self, name):
        path = HERE / name
        if not path.exists():
            build()
        self.lib = _c.CDLL(str(path))
        L = self.lib
        L.oracle_smix.argtypes, L.oracle_smix.restype = [U64], U64
        L.oracle_urand.argtypes, L.oracle_urand.restype = [U64, I, I, I], F
        L.oracle_set_threads.argtypes = [I]
        L.oracle_get_threads.restype = I
        bg = [P, P, I, P, P, P, P, P, I, F, F, F, F, I, I, I, P, I, P]
        L.oracle_build_grid_f64.argtypes = bg
        L.oracle_build_grid_f32.argtypes = bg
        bf = [P, P, I, P, P, I, P, I, F, F, F, F, I, I, I, P, P]
        L.oracle_build_grid_fit_f64.argtypes = bf
        L.oracle_build_grid_fit_f32.argtypes = bf
        L.oracle_inter_exact_fit_f64.argtypes = [P, P, I, I, I, P, P, P, I, I, P, P, I, P, I, P, P]
        L.oracle_eg_f64.argtypes = _LIG + _GRID + _SF + _REC + [I, P, I, P, P, P, I, P, P, P, P]
        L.oracle_eg_f32.argtypes = _LIG + _GRID + _SF + _REC + [I, P, I, P, P, P]
        L.oracle_mc.argtypes = ([I] + _LIG + _GRID + _SF + _REC + [I, P, P, I, P, P, I, P, P, P, P, I, I, I, F, F, F,
                                                                  U64, I, P])
        L.oracle_stage_a_f64.argtypes = [P, P, P, I, P, I, P, P, I] + _GRID + [I, F, LL, P, P]


_LIBS = {}


def lib(name="liboracle.so"):
    if name not in _LIBS:
        _LIBS[name] = _Lib(name)
    return _LIBS[name]


def set_threads(n, name="liboracle.so"):
    lib(name).lib.oracle_set_threads(int(n))


def smix(x):
    return int(lib().lib.oracle_smix(x))


def urand(seed, chain, step, slot):
    return float(lib().lib.oracle_urand(seed, chain, step, slot))


def build_grid(rec_xyz, rec_types, sfp_arrays, prm_grid, origin, spacing, n, lig_types, per_term=False, f64=True):
    """Oracle grid [T, nx, ny, nz] ([5, T, ...] if per_term). sfp_arrays = (radius f32, hyd, don, acc int32)."""
    nx, ny, nz = map(int, n)
    rx = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
    rt = np.ascontiguousarray(rec_types, np.int32)
    lt = np.ascontiguousarray(lig_types, np.int32)
    rad, hyd, don, acc = [np.ascontiguousarray(a, t) for a, t in zip(sfp_arrays, (np.float32, np.int32, np.int32,
                                                                                  np.int32))]
    prm = np.ascontiguousarray(prm_grid, np.float32)
    out = np.zeros(((5 if per_term else 1) * len(lt) * nx * ny * nz,), np.float64 if f64 else np.float32)
    fn = lib().lib.oracle_build_grid_f64 if f64 else lib().lib.oracle_build_grid_f32
    fn(_ptr(rx), _ptr(rt), len(rx), _ptr(rad), _ptr(hyd), _ptr(don), _ptr(acc), _ptr(lt), len(lt), float(origin[0]),
       float(origin[1]), float(origin[2]), float(spacing), nx, ny, nz, _ptr(prm), int(per_term), _ptr(out))
    shape = ((5, len(lt)) if per_term else (len(lt),)) + (nx, ny, nz)
    return out.reshape(shape)


def build_grid_fit(rec_xyz, rec_ext, rad_rec, rad_lig, W, prm12, origin, spacing, n, f64=True):
    """Oracle fitted-function grid [n_lt, nx, ny, nz] (src/adock/fit/gridfit.py:13-58). W [n_rt, n_lt, 6],
    prm12 = cutoff, g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good, hb_bad, ion_rmin, rep_cap."""
    nx, ny, nz = (int(n),) * 3 if np.isscalar(n) else map(int, n)
    rx = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
    rt = np.ascontiguousarray(rec_ext, np.int32)
    rr = np.ascontiguousarray(rad_rec, np.float32)
    rl = np.ascontiguousarray(rad_lig, np.float32)
    Wc = np.ascontiguousarray(W, np.float32)
    assert Wc.ndim == 3 and Wc.shape[1] == len(rl) and Wc.shape[2] == 6
    pr = np.ascontiguousarray(prm12, np.float32)
    assert pr.shape == (12,)
    out = np.zeros((len(rl) * nx * ny * nz,), np.float64 if f64 else np.float32)
    fn = lib().lib.oracle_build_grid_fit_f64 if f64 else lib().lib.oracle_build_grid_fit_f32
    fn(_ptr(rx), _ptr(rt), len(rx), _ptr(rr), _ptr(rl), len(rl), _ptr(Wc), Wc.shape[0], float(origin[0]),
       float(origin[1]), float(origin[2]), float(spacing), nx, ny, nz, _ptr(pr), _ptr(out))
    return out.reshape(len(rl), nx, ny, nz)


def inter_exact_fit(X, lig_type, rec_idx, rec_xyz, rec_ext, rad_rec, rad_lig, W, prm12):
    """Oracle exact pairwise fitted inter energy [L, M] in FP64 (src/adock/search.py:73-89 with FittedPotential.energy,
    src/adock/fit/terms.py:339-344). X [L, M, A, 3], lig_type [L, A] (outside [0, n_lt): skipped), rec_idx [L],
    rec_xyz [R, N, 3], rec_ext [R, N] (outside [0, n_rt): skipped), W [n_rt, n_lt, 6], prm12 as build_grid_fit."""
    c = np.ascontiguousarray
    X = c(X, np.float32)
    L, M, A = X.shape[:3]
    lt = c(lig_type, np.int32).reshape(L, A)
    ri = c(rec_idx, np.int32).reshape(L)
    rx = c(rec_xyz, np.float32)
    re = c(rec_ext, np.int32)
    R, N = re.shape
    assert rx.shape == (R, N, 3)
    Wc = c(W, np.float32)
    rl = c(rad_lig, np.float32)
    assert Wc.ndim == 3 and Wc.shape[1] == len(rl) and Wc.shape[2] == 6
    pr = c(prm12, np.float32)
    assert pr.shape == (12,)
    out = np.zeros((L, M), np.float64)
    lib().lib.oracle_inter_exact_fit_f64(_ptr(X), _ptr(lt), L, M, A, _ptr(ri), _ptr(rx), _ptr(re), R, N,
                                         _ptr(c(rad_rec, np.float32)), _ptr(rl), len(rl), _ptr(Wc), Wc.shape[0],
                                         _ptr(pr), _ptr(out))
    return out


class OracleProblem:
    """Oracle twin of the product KernelProblem: takes the same packed numpy arrays.

    lig: dict from adock_amd.ligand.pack_ligands; grid: dict(flat, origin, lo, hi, rec_idx, offs, dims, spacing,
    penalty); sf: dict(prm (search order), rad, hyd, don, acc); rec: optional dict(rxyz [R,N,3], rtyp [R,N], rn [R])."""

    def __init__(self, AM, KM, lig, grid, sf, rec=None):
        self.AM, self.KM = AM, KM
        self.SD, self.DM = 7 + KM, 6 + KM
        c = np.ascontiguousarray
        self.lig = {k: (c(v) if isinstance(v, np.ndarray) else v) for k, v in lig.items()}
        self.grid = {k: (c(v) if isinstance(v, np.ndarray) else v) for k, v in grid.items()}
        self.sf = {k: c(v) for k, v in sf.items()}
        if rec is None:
            rec = {"rxyz": np.zeros((1, 1, 3), np.float32), "rtyp": np.zeros((1, 1), np.int32),
                   "rn": np.zeros(1, np.int32)}
        self.rec = {k: c(v) for k, v in rec.items()}

    def _common(self):
        L, G, S, R = self.lig, self.grid, self.sf, self.rec
        return ([_ptr(L[k]) for k in ("ref", "typ", "tm", "ax", "axp", "axu", "pr", "na", "nt", "np", "rg")]
                + [int(L["P"]), _ptr(L["gslot"]), self.AM, self.KM]
                + [_ptr(G[k]) for k in ("flat", "origin", "lo", "hi", "rec_idx", "offs", "dims")]
                + [float(G["spacing"]), float(G["penalty"])]
                + [_ptr(S[k]) for k in ("prm", "rad", "hyd", "don", "acc")]
                + [_ptr(R["rxyz"]), _ptr(R["rtyp"]), _ptr(R["rn"]), int(R["rxyz"].shape[1])])

    def energy_grad(self, chain_lig, state, exact=False, nbr=None, nbr_n=None, f64=True):
        """-> (e [N], g [N, DM], X [N, AM, 3], margin [N]) in FP64 (f64=True) or (e, g) in FP32."""
        cl = np.ascontiguousarray(chain_lig, np.int32)
        st = np.ascontiguousarray(state, np.float32)
        n = len(cl)
        if f64:
            e = np.zeros(n)
            g = np.zeros((n, self.DM))
            X = np.zeros((n, self.AM, 3))
            m = np.zeros(n)
            nb = np.ascontiguousarray(nbr, np.int32) if nbr is not None else None
            nbn = np.ascontiguousarray(nbr_n, np.int32) if nbr_n is not None else None
            lib().lib.oracle_eg_f64(*self._common(), int(exact), _ptr(cl), n, _ptr(st), _ptr(nb), _ptr(nbn),
                                    int(nb.shape[1]) if nb is not None else 0, _ptr(e), _ptr(g), _ptr(X), _ptr(m))
            return e, g, X, m
        e = np.zeros(n, np.float32)
        g = np.zeros((n, self.DM), np.float32)
        lib().lib.oracle_eg_f32(*self._common(), int(exact), _ptr(cl), n, _ptr(st), _ptr(e), _ptr(g))
        return e, g

    def run_mc(self, chain_lig, chain_id, cur, cur_e, best, best_e, init_opt, mc_steps, bfgs_iters, temp, amp_pos,
               amp_rot, seed, step_offset, exact=False, nbr=None, nbr_n=None, f64=True, libname="liboracle.so"):
        """In place on float32 numpy cur/cur_e/best/best_e (reference semantics). Returns per-chain eval counts."""
        for a in (cur, cur_e, best, best_e):
            assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
        cl = np.ascontiguousarray(chain_lig, np.int32)
        ci = np.ascontiguousarray(chain_id, np.int32)
        nb = np.ascontiguousarray(nbr, np.int32) if nbr is not None else None
        nbn = np.ascontiguousarray(nbr_n, np.int32) if nbr_n is not None else None
        ev = np.zeros(len(cl), np.int64)
        lib(libname).lib.oracle_mc(int(f64), *self._common(), int(exact), _ptr(cl), _ptr(ci), len(cl), _ptr(nb),
                                   _ptr(nbn), int(nb.shape[1]) if nb is not None else 0, _ptr(cur), _ptr(cur_e),
                                   _ptr(best), _ptr(best_e), int(init_opt), int(mc_steps), int(bfgs_iters),
                                   float(temp), float(amp_pos), float(amp_rot), int(seed), int(step_offset), _ptr(ev))
        return ev

    def stage_a(self, conf, quats, trans, ntrans, cap):
        conf = np.ascontiguousarray(conf, np.float32)
        quats = np.ascontiguousarray(quats, np.float32)
        trans = np.ascontiguousarray(trans, np.float32)
        ntrans = np.ascontiguousarray(ntrans, np.int32)
        L, C = conf.shape[:2]
        Q, NT = quats.shape[0], trans.shape[1]
        be = np.zeros(L * C * Q)
        bi = np.zeros(L * C * Q, np.int32)
        G = self.grid
        lib().lib.oracle_stage_a_f64(_ptr(conf), _ptr(self.lig["gslot"]), _ptr(self.lig["na"]), C, _ptr(quats), Q,
                                     _ptr(trans), _ptr(ntrans), NT,
                                     *[_ptr(G[k]) for k in ("flat", "origin", "lo", "hi", "rec_idx", "offs", "dims")],
                                     float(G["spacing"]), float(G["penalty"]), self.AM, float(cap), L * C * Q,
                                     _ptr(be), _ptr(bi))
        return be.reshape(L, C * Q), bi.reshape(L, C * Q)
