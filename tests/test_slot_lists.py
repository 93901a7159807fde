"""Host-built pair-slot lists (adock_amd.ligand.slot_lists): every intra pair appears once in the list of each of its
atoms, as the byte offsets of its slot (4 q) and of the other atom (16 p); lists ascend in q, start and length are
multiples of 8, the padding names the zero slot and the dummy atom, and the row ends with 8 padding entries. With the
atom count given (one atom per lane), lanes past the atoms take the second half of the longest lists."""
import math

import numpy as np
import pytest

from adock_amd import synth
from adock_amd.ligand import inc_stride, list_lengths, pack_ligands, pad128, pad_entry, slot_lists


def _heavy_pairs(lig):
    hv = np.where(lig.heavy)[0]
    remap = -np.ones(lig.n, np.int64)
    remap[hv] = np.arange(len(hv))
    return remap[lig.pairs] if len(lig.pairs) else np.zeros((0, 2), np.int64)


def _decode(ist):
    ist = np.asarray(ist, np.int64)
    return ist & 0xffff, (ist >> 16) & 0xff, (ist >> 24) & 0x7f


def _lane_atom(ist, na):
    """The atom whose list each lane walks: its own, or (a helper lane past the atoms) its partner's."""
    s, c, p = _decode(ist)
    lane = np.arange(len(ist))
    return np.where((lane >= na) & (c > 0), p, lane)


@pytest.mark.parametrize("args", [(7, 12, 0), (9, 33, 1), (8, 64, 15), (3, 35, 8)])
@pytest.mark.parametrize("helpers", [False, True])
def test_slot_lists_cover_every_pair_twice(args, helpers):
    lig = synth.make_ligand(*args)
    na = int(lig.heavy.sum())
    pl = _heavy_pairs(lig)
    P_MAX = max(1, len(pl)) + 5
    inc, ist = slot_lists(pl, P_MAX, na=na if helpers else None)
    Z = pad_entry(P_MAX)
    assert Z == 4 * pad128(P_MAX) | (16 * 64) << 16
    assert inc.shape == (inc_stride(P_MAX),) and inc.dtype == np.uint32
    assert (inc[-8:] == Z).all()
    starts, counts, _ = _decode(ist)
    owner = _lane_atom(ist, na)
    seen = set()
    per_atom = {}
    for lane in range(64):
        s, c, a = int(starts[lane]), int(counts[lane]), int(owner[lane])
        assert s % 8 == 0 and c % 8 == 0 and s + c <= len(inc) - 8
        per_atom.setdefault(a, []).append((s, c))
    for a, segs in per_atom.items():
        segs.sort()
        # an atom's segments are contiguous: together they are its one padded list
        for (s0, c0), (s1, _) in zip(segs, segs[1:]):
            assert s1 == s0 + c0
        s, c = segs[0][0], sum(x[1] for x in segs)
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


def _best_longest_walk(groups, helpers):
    """Smallest achievable longest walk (in groups of 8) when at most `helpers` atoms may hand the second half of their
    list to one helper each: the least D with ceil(g / 2) <= D for every atom above D and at most `helpers` of them."""
    for D in range(0, max(groups, default=0) + 1):
        over = [g for g in groups if g > D]
        if len(over) <= helpers and all(math.ceil(g / 2) <= D for g in over):
            return D
    return 0


@pytest.mark.parametrize("args", [(7, 12, 0), (9, 33, 1), (3, 35, 8), (21, 48, 6), (5, 60, 12), (8, 64, 15)])
def test_helper_lanes_plan(args):
    lig = synth.make_ligand(*args)
    na = int(lig.heavy.sum())
    pl = _heavy_pairs(lig)
    P_MAX = max(1, len(pl))
    _, plain = slot_lists(pl, P_MAX)
    inc, ist = slot_lists(pl, P_MAX, na=na)
    s0, c0, p0 = _decode(plain)
    s, c, p = _decode(ist)
    lane = np.arange(64)
    assert (p0 == lane).all()
    # partners pair an atom with a helper lane past the atoms, both ways; a helper continues its atom's list
    for a in range(64):
        if p[a] != a:
            h, b = (p[a], a) if a < na else (a, p[a])
            assert b < na <= h and p[h] == b and p[b] == h
            assert s[h] == s[b] + c[b] and c[b] + c[h] == c0[b] and s[b] == s0[b]
            assert c[b] >= c[h] > 0
    np.testing.assert_array_equal(list_lengths(ist)[:na], c0[:na])
    # the longest walk is the least any one-helper-per-atom split of the padded lists could reach
    assert int(c.max()) // 8 == _best_longest_walk(list(c0[:na] // 8), 64 - na)
    if na == 64:
        np.testing.assert_array_equal(ist, plain)


def test_slot_lists_two_atoms_per_lane_and_limit():
    """AM 128: the dummy atom is 128 and no lane helps another; more pairs than 16-bit slot offsets can name are
    refused."""
    pl = _heavy_pairs(synth.make_ligand(820, 120, 20))
    P = len(pl)
    inc, ist = slot_lists(pl, P, 128, na=120)
    assert inc.shape == (inc_stride(P, 128),) and (inc[-8:] == (4 * pad128(P) | (16 * 128) << 16)).all()
    s, c, p = _decode(ist)
    assert (p == np.arange(128)).all()
    assert int(c.sum()) >= 2 * P
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
        inc, ist = slot_lists(pr, P, AM, na=int(out["na"][l]))
        np.testing.assert_array_equal(out["inc"][l], inc)
        np.testing.assert_array_equal(out["ist"][l], ist)
    # the dense rows hold whole lists: DEG is the longest full list, not the longest segment
    _, plain = slot_lists(out["pr"][1, :out["np"][1]], P, AM)
    assert out["DEG"] == max(int(list_lengths(out["ist"][l]).max()) for l in range(2))
    assert out["DEG"] >= int(((plain >> 16) & 0xff).max())
