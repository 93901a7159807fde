"""The C-ABI library loads and exports every entry point declared in include/adock_hip.h; argument validation rejects
bad buckets before anything reaches the GPU (no compute calls here)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "adock_hip.h"
LIB = ROOT / "adock_amd" / "libadock_hip.so"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(adock_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for want in ("adock_build_grid", "adock_eg", "adock_mc", "adock_mc_counted", "adock_last_error"):
        assert want in names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.skip("libadock_hip.so not built (python -m adock_amd.build)")
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_bucket_validation_without_gpu():
    if not LIB.exists():
        pytest.skip("libadock_hip.so not built")
    from adock_amd import _lib
    L = _lib.lib()
    null = [None] * 200
    # AM = 32 is not a compiled bucket: rejected before any HIP call
    args = [32, 8] + null[:11] + [1, None, None, None, 0] + null[:7] + [1, 0.375, 10.0] + null[:5] + [None, None, None, 0, 0,
                                                                                       None, 0, None, None, None, None]
    st = L.adock_eg(*args)
    assert st == -2
    assert b"AM" in L.adock_last_error()
    args[0], args[1] = 64, 12
    assert L.adock_eg(*args) == -2
    assert L.adock_set_chains_per_block(-1) == -1
    assert L.adock_set_chains_per_block(5) == -1
    assert L.adock_set_line_search_width(3) == -1
    assert L.adock_set_line_search_width(0) == 0
    assert L.adock_get_exact_staging() == 2
    assert L.adock_set_exact_staging(3) == -1 and L.adock_set_exact_staging(-1) == -1
    assert L.adock_set_exact_staging(1) == 0 and L.adock_get_exact_staging() == 1
    assert L.adock_set_exact_staging(2) == 0
    assert L.adock_version().startswith(b"adock_hip")
    assert L.adock_mc_resident_waves(64, 12, 64, 1) == -2 and L.adock_mc_resident_waves(64, 16, -1, 0) < 0
    # buckets: AM 96 / 128 and KM up to 56 are accepted past the bucket check (then stopped at the null pointers);
    # AM 160 and KM 64 are refused
    for am, km, want in ((96, 16, -1), (128, 56, -1), (160, 8, -2), (64, 64, -2), (128, 60, -2)):
        args[0], args[1] = am, km
        assert L.adock_eg(*args) == want, (am, km)
    # AM 64 holds at most 2,016 distinct pairs: P_MAX up to 4,096 passes the bucket check, above it is refused
    args[0], args[1], args[13] = 64, 16, 4096
    assert L.adock_eg(*args) == -1
    args[13] = 4097
    assert L.adock_eg(*args) == -3 and b"P_MAX" in L.adock_last_error()
    args[13] = 1
    # fitted grid builder: null pointers and more receptor ext types than the LDS table holds are rejected,
    # empty grids are a no-op
    fit = [None, None, 1, None, None, 1, None, 16, 0.0, 0.0, 0.0, 0.375, 4, 4, 4, None, None, None]
    assert L.adock_build_grid_fit(*fit) == -1          # null pointers first
    assert L.adock_build_grid_fit(*(fit[:12] + [0, 4, 4] + fit[15:])) == 0
    buf = ctypes.create_string_buffer(64)
    host = ctypes.cast(buf, ctypes.c_void_p).value       # never dereferenced: the call fails its checks first
    bad_rt = [None, None, 0, None, host, 1, None, 300] + fit[8:15] + [host, host, None]
    assert L.adock_build_grid_fit(*bad_rt) == -4
    # fitted grids of several receptor sets: no sets is a no-op, a missing origin array or too many sets is rejected
    fs = [host, host, 2, 4, host, host, 1, host, 16, host, 0.375, 4, 4, 4, host, host, None]
    assert L.adock_build_grid_fit_sets(*(fs[:2] + [0] + fs[3:])) == 0
    assert L.adock_build_grid_fit_sets(*(fs[:9] + [None] + fs[10:])) == -1
    assert L.adock_build_grid_fit_sets(*(fs[:2] + [70000] + fs[3:])) == -1
    assert L.adock_build_grid_fit_sets(*(fs[:8] + [300] + fs[9:])) == -4
    # exact fitted inter energy: empty pose sets are a no-op; null pointers, negative sizes, no receptor set and
    # receptor atoms without their tables are rejected before any launch
    ex = [host, host, 2, 3, 5, host, None, None, 1, 0, None, None, 0, None, 0, host, host, None]
    assert L.adock_inter_exact_fit(*(ex[:2] + [0] + ex[3:])) == 0
    assert L.adock_inter_exact_fit(*(ex[:16] + [None, None])) == -1
    assert L.adock_inter_exact_fit(*(ex[:2] + [-1] + ex[3:])) == -1
    assert L.adock_inter_exact_fit(*(ex[:8] + [0] + ex[9:])) == -1
    assert L.adock_inter_exact_fit(*(ex[:9] + [7] + ex[10:])) == -1


def test_facade_rejects_mismatched_chain_buffers():
    """KernelProblem.run_mc / BatchRunner check every buffer's extent before anything reaches the kernels: a state
    buffer passed as an energy vector (or the reverse) is an AdockError, not an out-of-bounds write."""
    import numpy as np
    import torch
    from adock_amd import _lib
    from adock_amd.kernel import KernelProblem
    kp = KernelProblem.__new__(KernelProblem)
    kp.SD, kp.device = 15, torch.device("cpu")
    N = 4
    st, e = torch.zeros(N, 15), torch.zeros(N)
    kp.check_chain_buffers(N, st, e, st.clone(), e.clone())
    for bad in ((st, st.clone(), e, e.clone()), (st, e, e.clone(), e.clone()), (st[:, :8].contiguous(), e, st, e),
                (st, e.double(), st, e), (st, e, st, torch.zeros(N + 1))):
        with pytest.raises(_lib.AdockError):
            kp.check_chain_buffers(N, *bad)


def test_product_path_fails_loudly_without_gpu_or_library(monkeypatch, tmp_path):
    """No CPU fallback: without a GPU the facades raise, and a missing library is an error naming the build step."""
    import numpy as np
    import torch
    from adock_amd import _lib
    from adock_amd.config import sf_params
    from adock_amd.grid import build_grid
    from adock_amd.kernel import KernelProblem
    from helpers import CFG
    if not torch.cuda.is_available():
        with pytest.raises(_lib.AdockError):
            _lib.require_gpu()
        with pytest.raises(_lib.AdockError):
            KernelProblem(object())
        with pytest.raises(_lib.AdockError):
            build_grid(np.zeros((1, 3), np.float32), np.zeros(1, np.int32), sf_params(CFG, "vina"), np.zeros(3),
                       0.375, 4, device="cuda")
    with pytest.raises(_lib.AdockError):
        build_grid(np.zeros((1, 3), np.float32), np.zeros(1, np.int32), sf_params(CFG, "vina"), np.zeros(3), 0.375,
                   4, device="cpu")
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "missing.so")
    with pytest.raises(_lib.AdockError, match="not built"):
        _lib.lib()


def test_learned_and_batch_argument_validation_without_gpu():
    """adock_mc_learned / adock_mc_step / adock_mc_batch reject bad network modes, temperatures, p_plain, buckets and
    LDS layouts before any launch (host addresses stand in for device buffers: the calls fail their checks first)."""
    if not LIB.exists():
        pytest.skip("libadock_hip.so not built")
    from adock_amd import _lib
    L = _lib.lib()
    buf = ctypes.create_string_buffer(64)
    h = ctypes.cast(buf, ctypes.c_void_p).value
    lig = [64, 8] + [h] * 11 + [128, h, h, h, 8]
    grid = [h] * 7 + [1, 0.375, 10.0]
    sf = [h] * 5
    rec = [h, h, h, 16]
    chains = [h, h, 1, h, h, h, h]            # chain_lig, chain_id, n_chains, cur, cur_e, best, best_e

    def learned(**kw):
        a = dict(init_opt=1, mc_steps=2, step_offset=0, bfgs_iters=0, temp=1.2, amp_pos=2.0, amp_rot=2.0, seed=1,
                 w=[h] * 6, feat=h, log_buf=None, log_count=h, log_cap=0, use_net=1, p_plain=0.25, sig=(0.5, 0.15, 0.3),
                 net_mode=0, tau=0.1, log_all=0, am=64)
        a.update(kw)
        args = ([a["am"]] + lig[1:] + grid + sf + rec + chains +
                [a["init_opt"], a["mc_steps"], a["step_offset"], a["bfgs_iters"], a["temp"], a["amp_pos"], a["amp_rot"],
                 a["seed"]] + a["w"] + [a["feat"], a["log_buf"], a["log_count"], a["log_cap"], a["use_net"],
                                        a["p_plain"], *a["sig"], a["net_mode"], a["tau"], a["log_all"], None, None, None,
                                        None])
        return L.adock_mc_learned(*args)

    def rejected(status, code, words):
        return status == code and words in L.adock_last_error()

    assert rejected(learned(w=[None] * 6), -1, b"weights")      # network moves without weights
    assert rejected(learned(p_plain=1.5), -1, b"p_plain")
    assert rejected(learned(net_mode=2), -1, b"net_mode")
    assert rejected(learned(net_mode=1, tau=0.0), -1, b"tau")
    assert rejected(learned(am=96), -2, b"64 heavy atoms")     # two atoms per lane: not in the learned kernels
    assert rejected(learned(mc_steps=-1), -1, b"mc_steps")
    step = ([64] + lig[1:] + grid + sf + rec + chains +
            [0, 0, 0, 1.2, 2.0, 2.0, 1, h, h, None, h, 0, 1, 0.25, 0.5, 0.15, 0.3, 3, 0.1, 0, None, None])
    assert rejected(L.adock_mc_step(*step), -1, b"net_mode")
    batch = ([64] + lig[1:] + grid + sf + rec + [0, h, h, 1, None, None, 0, h, h, h, h, 1, 2, 0, 1.2, 2.0, 2.0, 1, 0,
                                                 None, None, 1, h, None, None, None, 0, 0, 2, None, None, None])
    assert L.adock_mc_batch(*batch) == -1               # lds_layout 2
    assert b"lds_layout" in L.adock_last_error()
