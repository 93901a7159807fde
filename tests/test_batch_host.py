"""Host-side batching logic: HBM admission and the chain cost model."""
import numpy as np

from adock_amd import synth
from adock_amd.batch import admit, chain_cost, queue_partition


def test_admission_respects_budget_and_covers_all():
    rng = np.random.default_rng(0)
    sizes = rng.uniform(2e6, 20e6, 64)
    for budget in (25e6, 60e6, 1e9):
        groups = admit(sizes, budget)
        flat = sorted(i for g in groups for i in g)
        assert flat == list(range(64))
        for g in groups:
            assert sizes[g].sum() <= budget or len(g) == 1
    assert len(admit(sizes, 1e12)) == 1
    assert len(admit([5.0, 5.0], 4.0)) == 2


def test_cost_grows_with_ligand_size():
    small = synth.make_ligand(1, 10, 2)
    big = synth.make_ligand(2, 55, 12)
    c = chain_cost([small, big])
    assert c[1] > 5 * c[0]


def _check_partition(cost, group, nq):
    order, start = queue_partition(cost, group, nq)
    assert sorted(order.tolist()) == list(range(len(cost)))
    assert start[0] == 0 and start[-1] == len(cost) and np.all(np.diff(start) >= 0) and len(start) == max(nq, 1) + 1
    for q in range(len(start) - 1 if nq > 1 else 0):
        o = order[start[q]:start[q + 1]]
        g = group[o]
        # each group's chains form one run inside a queue (one receptor at a time), runs in descending total cost,
        # longest chains first inside a run
        runs = np.split(np.arange(len(o)), np.nonzero(np.diff(g))[0] + 1) if len(o) else []
        assert len({int(g[r[0]]) for r in runs}) == len(runs)
        tot = [cost[o[r]].sum() for r in runs]
        assert all(a >= b for a, b in zip(tot, tot[1:]))
        for r in runs:
            assert np.all(np.diff(cost[o[r]]) <= 0)
    return order, start


def test_queue_partition_keeps_receptors_whole_and_balances():
    """64 receptor sets x 32 ligands x 32 chains over 8 queues (one per XCD): every chain exactly once, each receptor's
    chains in one queue, queue costs within one group's cost of each other; a single receptor is dealt over all
    queues; one queue is the longest-first order."""
    rng = np.random.default_rng(0)
    lig_cost = rng.uniform(1.0, 30.0, 64 * 32)
    cl = np.repeat(np.arange(64 * 32), 32)
    cost, group = lig_cost[cl], cl // 32
    order, start = _check_partition(cost, group, 8)
    qof = np.empty(len(cost), int)
    for q in range(8):
        qof[order[start[q]:start[q + 1]]] = q
    for g in range(64):
        assert len(np.unique(qof[group == g])) == 1
    loads = np.array([cost[order[start[q]:start[q + 1]]].sum() for q in range(8)])
    gmax = max(cost[group == g].sum() for g in range(64))
    assert loads.max() - loads.min() <= gmax
    one = np.zeros(len(cost), int)
    order1, start1 = _check_partition(cost, one, 8)
    loads1 = np.array([cost[order1[start1[q]:start1[q + 1]]].sum() for q in range(8)])
    assert loads1.min() > 0.9 * loads1.max()
    o, st = _check_partition(cost, group, 1)
    assert np.array_equal(o, np.argsort(-cost, kind="stable"))
    o, st = _check_partition(cost[:5], group[:5], 8)           # fewer chains than queues
    assert (np.diff(st) > 0).sum() <= 5
