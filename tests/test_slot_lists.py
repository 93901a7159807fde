"""Host-built pair-slot lists (adock_amd.ligand.slot_lists): every intra pair appears once in the list of each of its
atoms, as the byte offsets of its slot (4 q) and of the other atom (16 p); lists ascend in q, start and length are
multiples of 8, the padding names the zero slot and the dummy atom, and the row ends with 8 padding entries."""
import numpy as np
import pytest

from adock_amd import synth
from adock_amd.ligand import inc_stride, pack_ligands, pad128, pad_entry, slot_lists


def _heavy_pairs(lig):
    hv = np.where(lig.heavy)[0]
    remap = -np.ones(lig.n, np.int64)
    remap[hv] = np.arange(len(hv))
    return remap[lig.pairs] if len(lig.pairs) else np.zeros((0, 2), np.int64)


@pytest.mark.parametrize("args", [(7, 12, 0), (9, 33, 1), (8, 64, 15), (3, 35, 8)])
def test_slot_lists_cover_every_pair_twice(args):
    pl = _heavy_pairs(synth.make_ligand(*args))
    P_MAX = max(1, len(pl)) + 5
    inc, ist = slot_lists(pl, P_MAX)
    Z = pad_entry(P_MAX)
    assert Z == 4 * pad128(P_MAX) | (16 * 64) << 16
    assert inc.shape == (inc_stride(P_MAX),) and inc.dtype == np.uint32
    assert (inc[-8:] == Z).all()
    seen = set()
    for a in range(64):
        s, c = int(ist[a]) & 0xffff, int(ist[a]) >> 16
        assert s % 8 == 0 and c % 8 == 0 and s + c <= len(inc) - 8
        ent = inc[s:s + c].astype(np.int64)
        real = ent[ent != Z]
        assert (ent[len(real):] == Z).all()
        q, px = (real & 0xffff), real >> 16
        assert (q % 4 == 0).all() and (px % 16 == 0).all()
        q, p = q // 4, px // 16
        assert (np.diff(q) > 0).all()
        for qq, pp in zip(q, p):
            side = 0 if pl[qq, 0] == a else 1
            assert pl[qq, side] == a and pl[qq, 1 - side] == pp and pp != a
            seen.add((int(qq), side))
    assert len(seen) == 2 * len(pl)


def test_slot_lists_two_atoms_per_lane_and_limit():
    """AM 128: the dummy atom is 128; more pairs than 16-bit slot offsets can name are refused."""
    pl = _heavy_pairs(synth.make_ligand(820, 120, 20))
    P = len(pl)
    inc, ist = slot_lists(pl, P, 128)
    assert inc.shape == (inc_stride(P, 128),) and (inc[-8:] == (4 * pad128(P) | (16 * 128) << 16)).all()
    assert int((ist >> 16).sum()) >= 2 * P
    with pytest.raises(ValueError):
        slot_lists(pl[:1], 16300, 128)


def test_pack_ligands_rows_match_per_ligand_lists():
    preps = [synth.make_ligand(7, 12, 0), synth.make_ligand(9, 33, 1)]
    slot = np.arange(32)[None, :]
    AM, KM, out = pack_ligands(preps, slot, [0, 0])
    P = int(out["P"])
    assert out["inc"].shape == (2, inc_stride(P))
    for l, p in enumerate(preps):
        pr = out["pr"][l, :out["np"][l]]
        # the packed pair list is the ligand's (heavy-atom) pair set, reordered by term class
        assert sorted(map(tuple, pr.tolist())) == sorted(map(tuple, _heavy_pairs(p).tolist()))
        inc, ist = slot_lists(pr, P, AM)
        np.testing.assert_array_equal(out["inc"][l], inc)
        np.testing.assert_array_equal(out["ist"][l], ist)
    assert out["DEG"] == max(int((out["ist"][l] >> 16).max()) for l in range(2))
