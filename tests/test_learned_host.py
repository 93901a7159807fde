"""Host-side training-set builders of the learned proposal (CPU tensors)."""
import torch

from adock_amd.learned import entity_dataset, move_weights


def test_entity_dataset_targets_the_moved_entity_only():
    X = torch.arange(5 * 128, dtype=torch.float32).reshape(5, 128)
    gain = torch.tensor([0.5, -1.0, 7.0, 0.2, 1.0])
    entity = torch.tensor([0, 1, 3, -1, 20])   # translation, rotation, torsion 1, a displacement, torsion 18
    Xo, T, W = entity_dataset(X, gain, entity)
    assert Xo.shape == (3, 128) and torch.equal(Xo, X[:3])
    assert T.shape == W.shape == (3, 23)
    assert torch.equal(W.sum(1), torch.ones(3))
    assert W[0, 0] == 1 and W[1, 3] == 1 and W[2, 8] == 1
    assert T[0, 0] == 0.5 and T[1, 3] == 0.0 and T[2, 8] == 5.0   # clipped to [0, 5]
    assert float((T * (1 - W)).abs().sum()) == 0.0


def test_move_weights_modes():
    mask = torch.ones(3, 23)
    gain = torch.tensor([2.0, -0.5, 9.0])
    acc = move_weights(mask, gain, "accepted")
    imp = move_weights(mask, gain, "improving")
    assert torch.allclose(acc[:, 0], torch.tensor([3.0, 0.25, 6.0]))
    assert torch.allclose(imp[:, 0], torch.tensor([2.0, 0.0, 5.0]))


def test_entity_dataset_best_target():
    """With best_gap the target is the improvement of the chain's best: gain + (best - current), clipped at 0."""
    X = torch.zeros(3, 128)
    gain = torch.tensor([2.0, 0.5, 1.0])
    best_gap = torch.tensor([-1.0, -1.0, 0.0])      # current above best by 1, 1, 0
    _, T, W = entity_dataset(X, gain, torch.tensor([0, 1, 2]), best_gap=best_gap)
    assert T[0, 0] == 1.0 and T[1, 3] == 0.0 and T[2, 7] == 1.0


def test_time_to_gap_interpolates_in_log_log():
    """The matched-gap time: log-log interpolation between the last budget above the target and the first at or
    below it; exact at a budget; None when no budget reaches the target or the first one already does."""
    import math
    from adock_amd.experiments import _time_to_gap
    rows = [{"m_gap": 0.4, "m_s": 1.0}, {"m_gap": 0.1, "m_s": 4.0}, {"m_gap": 0.025, "m_s": 16.0}]
    # gap 0.2 lies halfway (in log) between 0.4 and 0.1, so the time is halfway (in log) between 1 and 4
    assert math.isclose(_time_to_gap(rows, "m", 0.2), 2.0, rel_tol=1e-12)
    assert math.isclose(_time_to_gap(rows, "m", 0.05), 8.0, rel_tol=1e-12)
    assert math.isclose(_time_to_gap(rows, "m", 0.1), 4.0, rel_tol=1e-12)
    assert _time_to_gap(rows, "m", 0.01) is None
    assert _time_to_gap(rows, "m", 0.5) is None
