"""What the trained proposal network gains over the plain random move, measured at reduced size (the full-size
experiment is `bench.py --workload learned`): the network is trained with the library's own kernels on moves logged by
the plain method on 16 training complexes, then plain search, learned search (one launch, inference inside the chain
loop) and a control with a zero-output network (network moves = the Gaussian noise alone) run on 16 disjoint complexes
at equal MC-step budgets, each scored by its mean best-score gap to a long plain run.

The test records the gaps, the learned search's time against the plain one's at equal steps and the speed-up at a
matched gap (printed, and returned in the experiment's rows). It asserts what must hold whatever the gain: the training
loss goes down, every gap is finite, and the learned search is not worse than plain by more than 0.5 kcal/mol at the
largest budget (the exact score still judges every move). At full size (512 unseen complexes x 64 chains) the learned
search reaches a 0.1 kcal/mol gap 0.98-1.08x as fast as plain, and the zero-output control 1.05x: the small gain comes
from the shape of the network move (all degrees of freedom at once, mixed with the plain move), not from what the
network learned."""
import math

import pytest

from adock_amd import experiments as ex

pytestmark = pytest.mark.gpu


def test_learned_proposal_gain_recorded():
    print("[gpu] learned proposal vs plain at equal MC-step budgets (16 train / 16 eval complexes)", flush=True)
    out = ex.learned(n_train=16, train_chains=16, train_steps=60, n_eval=16, eval_chains=16, budgets=(10, 30, 60),
                     full_steps=300, epochs=3, log=lambda m: print("  " + m, flush=True))
    losses = out["train_loss"]
    assert losses and losses[-1] < losses[0], losses
    rows = out["rows"]
    assert [r["steps"] for r in rows] == [10, 30, 60]
    for r in rows:
        assert math.isfinite(r["plain_gap"]) and math.isfinite(r["learned_gap"])
    last = rows[-1]
    assert last["learned_gap"] <= last["plain_gap"] + 0.5, last
    step_ratio = sum(r["learned_s"] for r in rows) / sum(r["plain_s"] for r in rows)
    print(f"  gaps (plain / learned): " +
          ", ".join(f"S={r['steps']}: {r['plain_gap']:.3f} / {r['learned_gap']:.3f}" for r in rows), flush=True)
    print(f"  learned search time / plain search time at equal steps: {step_ratio:.2f}; speed-up at a "
          f"{out['gap_target_kcal']} kcal/mol gap: {out['speedup_at_matched_gap']}", flush=True)
