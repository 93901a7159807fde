"""Host-side batching logic: HBM admission and the chain cost model."""
import numpy as np

from adock_amd import synth
from adock_amd.batch import admit, chain_cost


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
