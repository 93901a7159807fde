"""The product path has no CPU fallback: without libadock_hip.so every entry point raises, and the grid builder refuses
a CPU device instead of computing there."""
import numpy as np
import pytest

from adock_amd import _lib, workload
from adock_amd.config import sf_params


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "libadock_hip.so")
    with pytest.raises(_lib.AdockError, match="not built"):
        _lib.lib()


def test_grid_builder_refuses_cpu():
    from adock_amd.grid import build_grid
    sfp = sf_params(workload.CFG, "vina")
    xyz = np.zeros((4, 3), np.float32)
    types = np.zeros(4, np.int32)
    with pytest.raises(_lib.AdockError, match="GPU only"):
        build_grid(xyz, types, sfp, np.zeros(3), 0.375, (4, 4, 4), lig_types=[0], device="cpu")
