"""End-to-end docking controls on a synthetic complex with a known native pose (the reference's controls on 1STP,
tests/test_docking.py, restated: its fixture needs the PDB mirror and RDKit). The receptor is grown around the ligand
in its reference conformation (synth.make_native_complex), so the native pose is translation 0, identity rotation,
torsions 0.

positive: top-1 pose within 2 A of the native pose, box centred on it and shifted 3 A (axis and diagonal)
negative: box displaced 15 A -> no output pose within 2 A
local minimisation from the native pose stays within 0.5 A
the same positive / negative pair through the fitted function (v2_fitted)
"""
import numpy as np
import pytest
import torch

from adock_amd import synth, workload
from adock_amd.kernel import KernelProblem
from adock_amd.search import dock_problem_kernel
from helpers import CFG

pytestmark = pytest.mark.gpu
POS_RMSD, NEG_SHIFT, OFFSET, MIN_RMSD, BOX = 2.0, 15.0, 3.0, 0.5, 20.0


def _cfg():
    cfg = dict(CFG["search"]["v2_default"])
    cfg.update(n_chains=128, mc_steps=64, mc_segment=32, min_total_chains=0, polish_top=16, polish_iters=50, n_out=9)
    return cfg


def _rmsd(X, native, heavy):
    return np.sqrt(((X[:, heavy] - native[heavy]) ** 2).sum(-1).mean(-1))


@pytest.mark.parametrize("seed", [5, 11])
def test_redock_controls(seed):
    print(f"[gpu] docking controls, native complex seed {seed}", flush=True)
    cx = synth.make_native_complex(seed, 20, 4)
    boxes = {"centred": np.zeros(3), "offset3": np.array([OFFSET, 0.0, 0.0]),
             "offset3_diag": np.full(3, OFFSET / np.sqrt(3.0)), "displaced15": np.array([0.0, 0.0, NEG_SHIFT])}
    recs = [workload.receptor_dict(cx["receptor"], c, BOX) for c in boxes.values()]
    lig = cx["lig"]
    prob = workload.gpu_problem(recs, [lig] * len(recs), list(range(len(recs))))
    res = dock_problem_kernel(prob, list(boxes.values()), [np.full(3, BOX)] * len(recs), _cfg(), seed=0)
    heavy = lig.types >= 0
    out = {k: _rmsd(r["coords"][:, :lig.n], cx["native"], heavy) for k, r in zip(boxes, res)}
    print({k: (round(float(v[0]), 2), round(float(v.min()), 2)) for k, v in out.items()}, flush=True)
    assert out["centred"][0] < POS_RMSD
    assert out["offset3"][0] < POS_RMSD
    assert out["offset3_diag"][0] < POS_RMSD
    assert out["displaced15"].min() >= POS_RMSD


def test_native_pose_local_minimisation():
    print("[gpu] native pose local minimisation", flush=True)
    cx = synth.make_native_complex(5, 20, 4)
    rec = workload.receptor_dict(cx["receptor"], np.zeros(3), BOX)
    lig = cx["lig"]
    prob = workload.gpu_problem([rec], [lig], [0])
    kp = KernelProblem(prob)
    st = kp.pack(torch.zeros(1, 3, device="cuda"), torch.tensor([[1.0, 0.0, 0.0, 0.0]], device="cuda"),
                 torch.zeros(1, prob.lb.K, device="cuda"))
    cur, best = st.clone(), st.clone()
    ce, be = torch.zeros(1, device="cuda"), torch.full((1,), float("inf"), device="cuda")
    kp.run_mc(torch.zeros(1, dtype=torch.int32, device="cuda"), torch.zeros(1, dtype=torch.int32, device="cuda"),
              cur, ce, best, be, True, 0, 100, 1.2, 2.0, 2.0, 0, 0)
    from adock_amd.pose import pose_coords
    t, q, th = kp.unpack(best, prob.lb.K)
    X = pose_coords(prob.lb, t[None], q[None], torch.zeros_like(t)[None], th[None])[0, 0].cpu().numpy()
    r = _rmsd(X[None, :lig.n], cx["native"], lig.types >= 0)[0]
    e0 = kp.energy_grad(torch.zeros(1, dtype=torch.int32, device="cuda"), st)[0]
    print(f"  rmsd after minimisation {r:.3f} A, energy {float(e0):.3f} -> {float(be):.3f}", flush=True)
    assert float(be) <= float(e0) and r < MIN_RMSD


def test_fitted_preset_controls():
    """The fitted function on the kernel search (the reference's test_v2_fitted_preset_controls_1stp): centred box
    top-1 within 2 A, box displaced 15 A -> no pose within 2 A."""
    print("[gpu] fitted-function docking controls", flush=True)
    from adock_amd.fit.fdock import DATA, dock_fitted, load_fitted
    cx = synth.make_native_complex(5, 20, 4)
    spec, w = load_fitted(DATA / "adock_fitted.json", CFG)
    rec = workload.receptor_dict(cx["receptor"], np.zeros(3), BOX)
    boxes = {"centred": np.zeros(3), "displaced15": np.array([0.0, 0.0, NEG_SHIFT])}
    cxs = [synth.fitted_complex(rec, cx["lig"], c, BOX, spec, seed=0) for c in boxes.values()]
    cfg = {**CFG, "search": {**CFG["search"], "ctl": _cfg()}}
    res = dock_fitted(cxs, spec, w, mode="ctl", seed=0, cfg=cfg)
    heavy = cx["lig"].types >= 0
    out = {k: _rmsd(r["coords"], cx["native"], heavy) for k, r in zip(boxes, res)}
    print({k: (round(float(v[0]), 2), round(float(v.min()), 2)) for k, v in out.items()}, flush=True)
    assert out["centred"][0] < POS_RMSD
    assert out["displaced15"].min() >= POS_RMSD
