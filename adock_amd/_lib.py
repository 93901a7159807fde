"""ctypes binding of libadock_hip.so (C ABI in include/adock_hip.h).

The library is built in-tree by adock_amd.build. There is no fallback: if the library is missing or the GPU is not
visible, every entry point raises.
"""
import ctypes
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libadock_hip.so"

P, I, F, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_ulonglong
LIG = [I, I] + [P] * 11 + [I, P, P, P, I]   # AM, KM, ref..rg, P_MAX, gslot, inc, ist, DEG_MAX
GRID = [P] * 7 + [I, F, F]                  # grid..gdims, T, spacing, pen
SF = [P] * 5
REC = [P, P, P, I]


class AdockError(RuntimeError):
    pass


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise AdockError(f"{LIB_PATH} is not built: run `python -m adock_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(str(LIB_PATH))
        L.adock_version.restype = ctypes.c_char_p
        L.adock_last_error.restype = ctypes.c_char_p
        L.adock_build_grid.argtypes = [P, P, I, P, P, P, P, P, I, F, F, F, F, I, I, I, P, I, P, P]
        L.adock_build_grid_fit.argtypes = [P, P, I, P, P, I, P, I, F, F, F, F, I, I, I, P, P, P]
        L.adock_build_grid_fit_sets.argtypes = [P, P, I, I, P, P, I, P, I, P, F, I, I, I, P, P, P]
        L.adock_inter_exact_fit.argtypes = [P, P, I, I, I, P, P, P, I, I, P, P, I, P, I, P, P, P]
        L.adock_eg.argtypes = LIG + GRID + SF + REC + [I, P, I, P, P, P, P]
        L.adock_mc.argtypes = LIG + GRID + SF + REC + [I, P, P, I, P, P, I, P, P, P, P, I, I, I, F, F, F, U64, I, P]
        L.adock_mc_counted.argtypes = (LIG + GRID + SF + REC +
                                       [I, P, P, I, P, P, I, P, P, P, P, I, I, I, F, F, F, U64, I, P, P])
        L.adock_mc_batch.argtypes = (LIG + GRID + SF + REC +
                                     [I, P, P, I, P, P, I, P, P, P, P, I, I, I, F, F, F, U64, I, P, P, I, P, P, P, P, I,
                                      I, I, P, P, P])
        L.adock_mc_batch_noalloc.argtypes = (LIG + GRID + SF + REC +
                                             [I, P, P, I, P, P, I, P, P, P, P, I, I, I, F, F, F, U64, I, P, P, I, I, P,
                                              P, I, I, P, P, P])
        L.adock_batch_reset.argtypes = [P, I, P, I, P, P]
        LL = ctypes.c_longlong
        L.adock_mc_step.argtypes = (LIG + GRID + SF + REC +
                                    [P, P, I, P, P, P, P, I, I, I, F, F, F, U64, P, P, P, P, I, I, F, F, F, F, I, F, I,
                                     P, P])
        L.adock_mc_learned.argtypes = (LIG + GRID + SF + REC +
                                       [P, P, I, P, P, P, P, I, I, I, I, F, F, F, U64] + [P] * 6 +
                                       [P, P, P, I, I, F, F, F, F, I, F, I, P, P, P, P])
        L.adock_mlp_infer.argtypes = [I, P, P, P, P, P, P, P, P, P]
        L.adock_gemm_f32.argtypes = [I, I, I, I, I, F, P, I, P, I, F, P, I, P]
        L.adock_gemm_f32_splitk.argtypes = [I, I, I, I, I, F, P, I, P, I, F, P, I, I, P, P]
        L.adock_bias_act.argtypes = [I, I, P, P, I, P]
        L.adock_relu_bwd.argtypes = [LL, P, P, P]
        L.adock_colsum.argtypes = [I, I, P, P, P]
        L.adock_mse.argtypes = [LL, P, P, P, F, P, P, P]
        L.adock_adam.argtypes = [LL, P, P, P, P, F, F, F, F, I, P]
        L.adock_to_bf16.argtypes = [LL, P, P, P]
        L.adock_set_chains_per_block.argtypes = [I]
        L.adock_get_chains_per_block.restype = I
        L.adock_stage_a.argtypes = [I, P, P, P, I, I, P, I, P, P, I] + [P] * 7 + [F, F, F, P, P, P]
        L.adock_set_line_search_width.argtypes = [I]
        L.adock_get_line_search_width.restype = I
        L.adock_set_exact_staging.argtypes = [I]
        L.adock_get_exact_staging.restype = I
        L.adock_mc_resident_waves.argtypes = [I, I, I, I]
        L.adock_mc_resident_waves.restype = I
        for name in ("adock_build_grid", "adock_build_grid_fit", "adock_build_grid_fit_sets", "adock_inter_exact_fit", "adock_eg", "adock_mc", "adock_mc_counted", "adock_mc_batch",
                     "adock_mc_batch_noalloc", "adock_batch_reset", "adock_set_chains_per_block", "adock_set_line_search_width", "adock_set_exact_staging", "adock_mc_step",
                     "adock_mc_learned",
                     "adock_mlp_infer", "adock_gemm_f32", "adock_gemm_f32_splitk", "adock_stage_a",
                     "adock_bias_act", "adock_relu_bwd", "adock_colsum", "adock_mse", "adock_adam", "adock_to_bf16"):
            getattr(L, name).restype = I
        _LIB = L
    return _LIB


def check(status):
    if status != 0:
        raise AdockError(f"libadock_hip status {status}: {lib().adock_last_error().decode()}")


def ptr(t):
    """Device pointer of a CUDA (ROCm) tensor; None -> NULL."""
    if t is None:
        return None
    if not t.is_cuda:
        raise AdockError("adock_amd kernels take device tensors (torch cuda/ROCm)")
    return t.data_ptr()


def stream_of(device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_gpu():
    if not torch.cuda.is_available():
        raise AdockError("adock_amd needs a visible GPU (MI355X, gfx950)")
