"""Ligand preparation from a heavy-atom graph: rigid-fragment torsion tree, intra-ligand pair list, radius of gyration,
and the packed per-ligand layout the search kernels read.

Conventions follow the reference so that packed arrays are interchangeable with its `pack_host` output:
- fragments = connected components after removing rotatable bonds; root = fragment with the smallest maximum BFS
  depth, ties to the larger heavy-atom count, then the lowest fragment index (src/adock/ligand.py:40-97);
- torsion k: axis (a, b) from the parent-side atom to the child-side atom, mask = atoms of the child subtree;
  torsions ordered leaf-to-root by a stable sort on depth (src/adock/ligand.py:98-109);
- intra pairs: heavy atoms in different fragments and more than 3 bonds apart (src/adock/ligand.py:111-119);
- ref = coordinates minus the heavy-atom centroid; rg over heavy atoms (src/adock/ligand.py:27-38).
The reference builds the same from an RDKit molecule; here the graph comes straight from the synthetic generator.
"""
import math
from collections import deque

import numpy as np


def _components(n, bonds):
    """Connected components of n atoms under `bonds` as labels 0.. numbered by each component's lowest atom."""
    root = list(range(n))

    def find(a):
        while root[a] != a:
            root[a] = root[root[a]]
            a = root[a]
        return a

    for i, j in bonds:
        ri, rj = find(i), find(j)
        if ri != rj:
            root[max(ri, rj)] = min(ri, rj)     # the lower atom stays the representative
    rep = np.array([find(a) for a in range(n)], np.int64)
    _, labels = np.unique(rep, return_inverse=True)   # representatives ascend with their lowest atom
    return labels.astype(np.int64)


def _level_sweep(nbrs, start):
    """Breadth-first sweep of the fragment graph from `start`, one level at a time (a level's fragments in discovery
    order, each one's neighbours in adjacency order). Returns (discovery order, depth array (-1 = unreached), parent
    link per fragment: (parent, atom on the parent side, atom on this side), None at the start)."""
    nf = len(nbrs)
    depth = np.full(nf, -1, np.int64)
    up = [None] * nf
    depth[start] = 0
    order, level = [start], [start]
    while level:
        nxt = []
        for u in level:
            for v, a, b in nbrs[u]:
                if depth[v] < 0:
                    depth[v] = depth[u] + 1
                    up[v] = (u, a, b)
                    nxt.append(v)
        order += nxt
        level = nxt
    return order, depth, up


class LigandPrep:
    """Heavy-atom ligand with its torsion tree.

    xyz [n,3] float, types [n] int (XS index, -1 = hydrogen), bonds [(i, j)], rot_bonds [(i, j)] (i < j)."""

    def __init__(self, xyz, types, bonds, rot_bonds, n_tors_vina=None):
        xyz = np.asarray(xyz, np.float64)
        self.n = len(xyz)
        self.types = np.asarray(types, np.int64)
        self.heavy = self.types >= 0
        self.bonds = [(int(min(i, j)), int(max(i, j))) for i, j in bonds]
        self.rot_bonds = sorted((int(min(i, j)), int(max(i, j))) for i, j in rot_bonds)
        self.center = xyz[self.heavy].mean(0)
        self.ref = (xyz - self.center).astype(np.float32)
        self._build_tree()
        self._build_pairs()
        h = self.ref[self.heavy]
        self.rg = float(np.sqrt(((h - h.mean(0)) ** 2).sum(1).mean())) if len(h) > 1 else 1.0
        self.n_tors_vina = self.n_tors if n_tors_vina is None else n_tors_vina

    def _adjacency(self, skip=()):
        adj = [[] for _ in range(self.n)]
        skip = set(skip)
        for i, j in self.bonds:
            if (i, j) in skip:
                continue
            adj[i].append(j)
            adj[j].append(i)
        return adj

    def _build_tree(self):
        """Rigid fragments and the torsion tree. Fragments: connected components of the bond graph without the
        rotatable bonds, numbered by their lowest atom index. Root: the fragment whose farthest fragment (in bonds of
        the fragment tree) is nearest, ties to more heavy atoms, then the lower number. Torsion of a non-root fragment
        f: axis (atom on the parent side, atom in f) of the bond to its parent, moving f and everything below it;
        torsions leaf-to-root by depth, ties in breadth-first discovery order from the root (fragments adjacent in
        rotatable-bond order)."""
        n = self.n
        frag = _components(n, [b for b in self.bonds if b not in set(self.rot_bonds)])
        nf = int(frag.max()) + 1 if n else 0
        self.frag = frag
        # fragment adjacency, in rotatable-bond order: (neighbour fragment, atom here, atom there)
        nbrs = [[] for _ in range(nf)]
        for i, j in self.rot_bonds:
            nbrs[frag[i]].append((frag[j], i, j))
            nbrs[frag[j]].append((frag[i], j, i))
        heavy_per_frag = np.bincount(frag[self.heavy], minlength=nf)
        sweeps = [_level_sweep(nbrs, f) for f in range(nf)]
        root = min(range(nf), key=lambda f: (int(sweeps[f][1][sweeps[f][0]].max()), -int(heavy_per_frag[f]), f))
        order, depth, up = sweeps[root]
        if len(order) != nf:
            raise ValueError("ligand graph is disconnected")
        self.root = root
        # atoms moved by each fragment's torsion: its own atoms and its descendants', gathered leaf-up
        moved = np.zeros((nf, n), bool)
        moved[frag, np.arange(n)] = True
        for f in order[:0:-1]:
            moved[up[f][0]] |= moved[f]
        rank = {f: k for k, f in enumerate(order)}
        kids = sorted(order[1:], key=lambda f: (-depth[f], rank[f]))
        self.tors_axis = np.array([[up[f][1], up[f][2]] for f in kids], np.int64).reshape(-1, 2)
        self.tors_mask = moved[kids].reshape(-1, n)
        self.n_tors = len(kids)

    def topo_distance(self):
        adj = self._adjacency()
        D = np.full((self.n, self.n), np.inf)
        for s in range(self.n):
            D[s, s] = 0
            dq = deque([s])
            while dq:
                u = dq.popleft()
                for v in adj[u]:
                    if D[s, v] == np.inf:
                        D[s, v] = D[s, u] + 1
                        dq.append(v)
        return D

    def _build_pairs(self):
        dm = self.topo_distance()
        hv = np.where(self.heavy)[0]
        pairs = [(i, j) for ii, i in enumerate(hv) for j in hv[ii + 1:] if self.frag[i] != self.frag[j] and dm[i, j] > 3]
        self.pairs = np.array(pairs, dtype=np.int64).reshape(-1, 2)


def pad128(P):
    return (int(P) + 127) // 128 * 128


def inc_stride(P_MAX, AM=64):
    """Length of one ligand's row of slot lists: two entries per pair, up to 7 padding entries per atom (each atom's
    list is padded to a multiple of 8), and a trailing group of 8 padding entries (what lanes past their list read)."""
    return 2 * pad128(P_MAX) + 8 * AM + 8


def pad_entry(P_MAX, AM=64):
    """The padding entry of the slot lists: the always-zero slot pad128(P_MAX) and the dummy atom (64 per atom slot
    of a lane: 64 for AM <= 64, else 128)."""
    return (4 * pad128(P_MAX)) | ((16 * (64 if AM <= 64 else 128)) << 16)


def slot_lists(pairs, P_MAX, AM=64, na=None):
    """Per-atom pair-slot lists of one ligand for the kernels' atomic-free pair gradient. Atom a's list holds, for
    every intra pair q of a (ascending in q), the word fs | px << 16: fs = 4 q, the byte offset of the pair's slot
    (dE/dr)/r in the kernels' slot array, and px = 16 p, the byte offset of the pair's other atom p in the atom array,
    so that the atom adds slot x (X_a - X_p) with two address additions. Lists are padded to a multiple of 8 entries
    with pad_entry (the zero slot and the dummy atom); 8 more padding entries end the row.

    Helper lanes (one atom per lane, AM = 64, na atoms given): the lanes na..63 hold no atom, so each takes the second
    half of one of the longest lists (longest first, ties by atom index): atom a keeps the first ceil(g / 2) of its g
    groups of 8 entries, its helper the other floor(g / 2), and the kernels add the helper's partial sum to the atom's.
    A lane walks only its own segment, so the longest walk of the wave shrinks.

    Returns (inc [inc_stride] uint32, ist [AM] int32 = start | count << 16 | partner << 24): start and count
    (multiples of 8) of the lane's segment of the row; partner = the helper lane of a split atom, the atom a helper
    lane works for, and the lane itself otherwise."""
    z = pad128(P_MAX)
    if 4 * z > 0xffff:
        raise ValueError(f"P_MAX = {P_MAX}: slot byte offsets exceed 16 bits (at most 16,256 pairs per ligand)")
    inc = np.full(inc_stride(P_MAX, AM), pad_entry(P_MAX, AM), np.uint32)
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    q = np.arange(len(pairs))
    real = pairs[:, 0] != pairs[:, 1]
    atom = np.concatenate([pairs[real, 0], pairs[real, 1]])
    part = np.concatenate([pairs[real, 1], pairs[real, 0]])
    qq = np.concatenate([q[real], q[real]])
    order = np.lexsort((qq, atom))
    atom, val = atom[order], ((4 * qq) | ((16 * part) << 16))[order].astype(np.uint32)
    cnt = np.bincount(atom, minlength=AM)[:AM]
    pc = (cnt + 7) // 8 * 8
    if pc.max(initial=0) > 0xff:
        raise ValueError("an atom's slot list exceeds 255 entries")
    start = np.concatenate([[0], np.cumsum(pc)[:-1]])
    if start[-1] + pc[-1] > len(inc) - 8:
        raise ValueError("slot lists exceed their row")
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    for a in np.nonzero(cnt)[0]:
        inc[start[a]:start[a] + cnt[a]] = val[first[a]:first[a] + cnt[a]]
    seg_s, seg_c, partner = start.copy(), pc.copy(), np.arange(AM)
    if na is not None and AM == 64 and int(na) < 64:
        na = int(na)
        g = pc[:na] // 8
        for h, a in zip(range(na, 64), np.argsort(-g, kind="stable")):
            if g[a] < 2:
                break
            keep = 8 * ((g[a] + 1) // 2)
            seg_c[a] = keep
            seg_s[h], seg_c[h] = start[a] + keep, pc[a] - keep
            partner[a], partner[h] = h, a
    ist = (seg_s | (seg_c << 16) | (partner << 24)).astype(np.int32)
    return inc, ist


def list_lengths(ist):
    """Full (padded) slot-list length of each lane's atom from an ist row: the lane's segment plus its helper's."""
    ist = np.asarray(ist, np.int64)
    s, c, p = ist & 0xffff, (ist >> 16) & 0xff, (ist >> 24) & 0x7f
    lane = np.arange(len(ist))
    # a split atom's helper segment starts right after the atom's own
    helped = (p != lane) & (s[p] == s + c) & (c > 0)
    return np.where(helped, c + c[p], np.where((p != lane) & (s == s[p] + c[p]), 0, c))


def bucket(x, step):
    return max(step, step * math.ceil(x / step))


def pair_classes(pairs, types):
    """Term class of each heavy-atom pair: 2 = h-bond pair (donor-acceptor either way), 1 = hydrophobic pair,
    0 = neither (XS type flags of config/adock.json)."""
    from .config import load_config
    t = load_config()["_xs_table"]
    hyd, don, acc = (np.asarray(t[k], bool) for k in ("hydrophobic", "donor", "acceptor"))
    ti, tj = types[pairs[:, 0]], types[pairs[:, 1]]
    hb = (don[ti] & acc[tj]) | (acc[ti] & don[tj])
    return np.where(hb, 2, np.where(hyd[ti] & hyd[tj], 1, 0))


def pack_ligands(preps, slot, rec_idx, floor=(0, 0)):
    """Packed ligand arrays in the search-kernel layout (src/adock/mckernel.py:547-592).

    slot [R, n_types] int (XS type -> grid slot of receptor r, -1 = not built), rec_idx [L]. floor (AM, KM): smallest
    buckets to use (several batches packed to one state layout and one kernel bucket). Returns (AM, KM, dict of numpy
    arrays: ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P, gslot, inc, ist, DEG); inc / ist: slot_lists of each
    ligand, DEG: the longest (padded) list."""
    L = len(preps)
    nh = [int(p.heavy.sum()) for p in preps]
    top = max(max(nh), int(floor[0]))
    AM = 64 if top <= 64 else bucket(top, 32)
    KM = bucket(max(max(p.n_tors for p in preps), 1, int(floor[1])), 8)
    pair_lists = []
    for p in preps:
        hv = np.where(p.heavy)[0]
        remap = -np.ones(p.n, np.int64)
        remap[hv] = np.arange(len(hv))
        pl = remap[p.pairs] if len(p.pairs) else np.zeros((0, 2), np.int64)
        if len(pl):
            # h-bond pairs first, then hydrophobic, then the rest: the kernels skip those terms in rounds without them
            cls = pair_classes(p.pairs, p.types)
            pl = pl[np.argsort(-cls, kind="stable")]
        pair_lists.append(pl)
    P = max(1, max(len(x) for x in pair_lists))
    out = {
        "ref": np.zeros((L, AM, 3), np.float32), "typ": np.zeros((L, AM), np.int32),
        "tm": np.zeros((L, KM, AM), np.uint8), "ax": np.zeros((L, KM, 2), np.int32),
        "axp": np.zeros((L, KM, 3), np.float32), "axu": np.zeros((L, KM, 3), np.float32),
        "pr": np.zeros((L, P, 2), np.int32), "na": np.zeros(L, np.int32), "nt": np.zeros(L, np.int32),
        "np": np.zeros(L, np.int32), "rg": np.zeros(L, np.float32), "gslot": np.zeros((L, AM), np.int32),
        "inc": np.zeros((L, inc_stride(P, AM)), np.uint32), "ist": np.zeros((L, AM), np.int32),
    }
    slot = np.asarray(slot)
    rec_idx = np.asarray(rec_idx)
    for l, p in enumerate(preps):
        hv = np.where(p.heavy)[0]
        remap = -np.ones(p.n, np.int64)
        remap[hv] = np.arange(len(hv))
        out["na"][l], out["nt"][l], out["np"][l], out["rg"][l] = len(hv), p.n_tors, len(pair_lists[l]), p.rg
        out["ref"][l, :len(hv)] = p.ref[hv]
        out["typ"][l, :len(hv)] = p.types[hv]
        if p.n_tors:
            out["tm"][l, :p.n_tors, :len(hv)] = p.tors_mask[:, hv]
            out["ax"][l, :p.n_tors] = remap[p.tors_axis]
            if (out["ax"][l, :p.n_tors] < 0).any():
                raise ValueError("torsion axis atom is not a heavy atom")
            pa, pb = p.ref[p.tors_axis[:, 0]], p.ref[p.tors_axis[:, 1]]
            u = pb - pa
            out["axp"][l, :p.n_tors] = pa
            out["axu"][l, :p.n_tors] = u / np.maximum(np.linalg.norm(u, axis=1, keepdims=True), 1e-6)
        if len(pair_lists[l]):
            out["pr"][l, :len(pair_lists[l])] = pair_lists[l]
        out["inc"][l], out["ist"][l] = slot_lists(pair_lists[l], P, AM, na=len(hv))
        out["gslot"][l, :len(hv)] = slot[rec_idx[l], out["typ"][l, :len(hv)]]
        if (out["gslot"][l, :len(hv)] < 0).any():
            raise ValueError("a ligand atom type has no grid")
    out["P"] = np.int32(P)
    out["DEG"] = np.int32(max(int(list_lengths(r).max()) for r in out["ist"]) if L else 0)
    return AM, KM, out


def quat_to_mat(q):
    q = np.asarray(q, np.float64)
    q = q / np.linalg.norm(q, axis=-1, keepdims=True)
    w, x, y, z = np.moveaxis(q, -1, 0)
    return np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                     2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                     2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1).reshape(q.shape[:-1] + (3, 3))


def random_quat(rng, n):
    q = rng.standard_normal((n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
