"""The fitted scoring function: a Vina-family function extended with extra terms, linear in its weights.

E(pose) = sum_m w_m F_m(pose). Each inter-molecular feature F_m sums, over (ligand heavy atom, receptor atom) pairs
within the cutoff, one basis function of the pair's distance times an indicator (or, for the ionic term, the charge
product) of the two atoms' extended types; five intra-ligand features are the Vina terms over the ligand's pairs
(src/adock/fit/terms.py). Extended types fold the extra attributes into the type index so that every term remains a
function of (r, ligand type, receptor type) and fits per-ligand-type grids:

    ligand ext type   = xs + N_XS * charge_class                         charge_class 0 neutral, 1 cation, 2 anion
    receptor ext type = xs + N_XS * (buried + 2 * charge_class) [+ 6 N_XS if side-chain split and side-chain atom]

Term names, their order (the order of the weight vector in a fitted-function JSON file) and the type encoding are the
reference's, so its config/adock_fitted.json files load unchanged. Charge classes of ligand atoms come from formal
charges (RDKit in the reference, src/adock/fit/terms.py:165-194); here they are an input array.
"""
import numpy as np
import torch

# basis functions, by index
BASIS = ("gauss1", "gauss2", "repulsion", "hydrophobic", "hbond", "ionic")
B_G1, B_G2, B_REP, B_HYD, B_HB, B_ION = range(6)
N_BASIS = len(BASIS)

# inter-molecular terms in weight-vector order: (name, basis, pair class) (src/adock/fit/terms.py:25-39)
INTER_TERMS = [
    ("gauss1", B_G1, "all"),
    ("gauss2", B_G2, "all"),
    ("repulsion", B_REP, "all"),
    ("hydrophobic", B_HYD, "hyd"),
    ("hbond", B_HB, "hb"),
    ("hyd_buried", B_HYD, "hyd_bur"),
    ("hbond_buried", B_HB, "hb_bur"),
    ("hbond_ligdonor", B_HB, "hb_ligdon"),
    ("hbond_charged", B_HB, "hb_chg"),
    ("metal_acceptor", B_HB, "metal"),
    ("ionic", B_ION, "ion"),
    ("polar_apolar", B_HYD, "mismatch"),
    ("gauss1_buried", B_G1, "bur"),
]
INTRA_TERMS = ["intra_gauss1", "intra_gauss2", "intra_repulsion", "intra_hydrophobic", "intra_hbond"]
TERM_NAMES = [t[0] for t in INTER_TERMS] + INTRA_TERMS
N_INTER = len(INTER_TERMS)
BASE5 = ["gauss1", "gauss2", "repulsion", "hydrophobic", "hbond"]
# basis parameters a fitted function may override for the protein-ligand terms only
INTER_PRM_KEYS = ("g1_off", "g1_w", "g2_off", "g2_w", "rep_off", "hyd_good", "hyd_bad", "hb_good", "hb_bad")
CHARGE_SIGN = np.array([0, 1, -1])     # charge class -> sign


class Spec:
    """Fixed (non-linear) part of a fitted function and the per-ext-type attribute tables
    (src/adock/fit/terms.py:56-126).

    base: "vina" | "vinardo" (radii and basis parameters); buried_threshold: receptor atoms with at least this many
    heavy neighbours within buried_radius are buried; ion_rmin: the ionic basis clamps r below it; inter_prm: overrides
    of INTER_PRM_KEYS for the protein-ligand basis; rep_cap: cap on the repulsion basis (None: uncapped);
    sc_radius_scale / sc_radius_delta: side-chain receptor atoms (beyond C-beta) get their own type block with radius
    xs_radius * scale - delta (either option enables the split)."""

    def __init__(self, cfg, base="vina", buried_threshold=1e9, buried_radius=10.0, ion_rmin=2.0, inter_prm=None,
                 rep_cap=None, sc_radius_scale=None, sc_radius_delta=None):
        sfc = cfg["scoring"][base]
        tab = cfg["_xs_table"]
        names = cfg["_xs_names"]
        self.base, self.cfg = base, cfg
        self.buried_threshold = float(buried_threshold)
        self.buried_radius = float(buried_radius)
        self.ion_rmin = float(ion_rmin)
        self.cutoff = float(sfc["cutoff"])
        self.prm = {"g1_off": sfc["gauss1"]["offset"], "g1_w": sfc["gauss1"]["width"],
                    "g2_off": sfc["gauss2"]["offset"], "g2_w": sfc["gauss2"]["width"],
                    "rep_off": sfc["repulsion"]["offset"],
                    "hyd_good": sfc["hydrophobic"]["good"], "hyd_bad": sfc["hydrophobic"]["bad"],
                    "hb_good": sfc["hbond"]["good"], "hb_bad": sfc["hbond"]["bad"]}
        bad = set(inter_prm or {}) - set(INTER_PRM_KEYS)
        if bad:
            raise ValueError(f"inter_prm keys not overridable: {sorted(bad)}")
        self.iprm = {**self.prm, **{k: float(v) for k, v in (inter_prm or {}).items()}}
        self.rep_cap = None if rep_cap is None else float(rep_cap)
        self.sc_radius_scale = None if sc_radius_scale is None else float(sc_radius_scale)
        self.sc_radius_delta = None if sc_radius_delta is None else float(sc_radius_delta)
        self.sc_split = self.sc_radius_scale is not None or self.sc_radius_delta is not None
        self.base_weights = {k: sfc[k]["weight"] for k in BASE5}
        self.rot_weight = sfc["rot_weight"]
        X = self.nxs = len(names)
        self.xs_radius = np.asarray(tab[sfc["radius_column"]], np.float32)
        self.xs_hyd = np.asarray(tab["hydrophobic"]).astype(bool)
        self.xs_don = np.asarray(tab["donor"]).astype(bool)
        self.xs_acc = np.asarray(tab["acceptor"]).astype(bool)
        self.xs_polar = np.array([n.split("_")[0] in ("N", "O") for n in names])
        self.xs_metal = np.array([n == "Met_D" for n in names])
        # ligand ext types: 3 charge blocks of X
        self.lig_xs = np.arange(3 * X) % X
        self.lig_chg = CHARGE_SIGN[np.arange(3 * X) // X]
        # receptor ext types: blocks b = buried + 2 * charge_class (+ 6 for side chains)
        nb = 12 if self.sc_split else 6
        blk = np.arange(nb * X) // X
        self.rec_xs = np.arange(nb * X) % X
        self.rec_bur = blk % 2 == 1
        self.rec_chg = CHARGE_SIGN[(blk // 2) % 3]
        self.rec_sc = blk >= 6
        self.n_lt, self.n_rt = 3 * X, nb * X
        scale = 1.0 if self.sc_radius_scale is None else self.sc_radius_scale
        delta = 0.0 if self.sc_radius_delta is None else self.sc_radius_delta
        r = self.xs_radius[self.rec_xs].astype(np.float64)
        self.rec_radius = np.where(self.rec_sc, r * scale - delta, r).astype(np.float32)

    def nonlinear(self):
        """The "nonlinear" section of a fitted-function JSON file."""
        d = {"base": self.base, "buried_threshold": self.buried_threshold, "buried_radius": self.buried_radius,
             "ion_rmin": self.ion_rmin, "cutoff": self.cutoff, **self.prm}
        over = {k: v for k, v in self.iprm.items() if v != self.prm[k]}
        if over:
            d["inter_prm"] = over
        for k in ("rep_cap", "sc_radius_scale", "sc_radius_delta"):
            if getattr(self, k) is not None:
                d[k] = getattr(self, k)
        return d

    def prm_tuple(self):
        """Intra-ligand (base) basis parameters in PairPotential.prm() order."""
        p = self.prm
        return (p["g1_off"], p["g1_w"], p["g2_off"], p["g2_w"], p["rep_off"], p["hyd_good"], p["hyd_bad"],
                p["hb_good"], p["hb_bad"], self.cutoff)

    def grid_prm(self):
        """The 12 floats the fitted grid builder takes: cutoff, the inter basis parameters, ion_rmin, rep_cap."""
        p = self.iprm
        cap = np.float32(3.0e38) if self.rep_cap is None else self.rep_cap
        return np.array([self.cutoff, p["g1_off"], p["g1_w"], p["g2_off"], p["g2_w"], p["rep_off"], p["hyd_good"],
                         p["hyd_bad"], p["hb_good"], p["hb_bad"], self.ion_rmin, cap], np.float32)

    def pair_classes(self):
        """Pair-class indicators [n_rt, n_lt] (the ionic class holds the charge product) (src/adock/fit/terms.py
        :128-156)."""
        lx, rx = self.lig_xs, self.rec_xs
        both = lambda lig, rec: rec[:, None] & lig[None, :]
        l_hyd, l_don, l_acc = self.xs_hyd[lx], self.xs_don[lx], self.xs_acc[lx]
        r_hyd, r_don, r_acc = self.xs_hyd[rx], self.xs_don[rx], self.xs_acc[rx]
        l_pol, r_pol, r_met = self.xs_polar[lx], self.xs_polar[rx], self.xs_metal[rx]
        hyd = both(l_hyd, r_hyd)
        hb = both(l_don, r_acc) | both(l_acc, r_don)
        bur = np.repeat(self.rec_bur[:, None], self.n_lt, 1)
        charged = (self.rec_chg != 0)[:, None] | (self.lig_chg != 0)[None, :]
        return {
            "all": np.ones((self.n_rt, self.n_lt), bool), "hyd": hyd, "hb": hb, "bur": bur,
            "hyd_bur": hyd & bur, "hb_bur": hb & bur,
            "hb_ligdon": both(l_don, r_acc & ~r_met), "hb_chg": hb & charged, "metal": both(l_acc, r_met),
            "ion": (self.rec_chg[:, None] * self.lig_chg[None, :]).astype(np.float32),
            "mismatch": both(l_hyd, r_pol) | both(l_pol, r_hyd),
        }

    def indicator(self):
        """I [n_rt, n_lt, N_BASIS, N_INTER] float32: weight of term m on basis k for a type pair."""
        cls = self.pair_classes()
        I = np.zeros((self.n_rt, self.n_lt, N_BASIS, N_INTER), np.float32)
        for m, (_, k, c) in enumerate(INTER_TERMS):
            I[:, :, k, m] = cls[c]
        return I

    def pair_table(self, w_inter):
        """W [n_rt, n_lt, N_BASIS] = sum_m w_m I[..., m]: the per-type-pair basis weights."""
        w = np.asarray(w_inter, np.float64)
        return np.tensordot(self.indicator().astype(np.float64), w, axes=([3], [0])).astype(np.float32)


def default_weights(spec):
    """Weights that reproduce the base function: its five weights on the base inter and intra terms, 0 elsewhere."""
    w = np.zeros(len(TERM_NAMES), np.float32)
    for k in BASE5:
        w[TERM_NAMES.index(k)] = spec.base_weights[k]
        w[TERM_NAMES.index("intra_" + k)] = spec.base_weights[k]
    return w


def receptor_ext_types(ann, spec):
    """Receptor ext types from an annotation dict: rec_types (XS), buried_count, rec_chg (charge class), rec_sc
    (side-chain flag, used when the spec splits side chains) (src/adock/fit/terms.py:224-230)."""
    bur = (np.asarray(ann["buried_count"]) >= spec.buried_threshold).astype(np.int64)
    ext = np.asarray(ann["rec_types"], np.int64) + spec.nxs * (bur + 2 * np.asarray(ann["rec_chg"], np.int64))
    if spec.sc_split:
        ext = ext + 6 * spec.nxs * np.asarray(ann["rec_sc"]).astype(np.int64)
    return ext


def ligand_ext_types(xs_types, charge_class, nxs):
    """Ligand ext types [N] from XS types (-1 = hydrogen, kept) and charge classes (0 / 1 / 2)."""
    xs = np.asarray(xs_types, np.int64)
    return np.where(xs >= 0, xs + nxs * np.asarray(charge_class, np.int64), -1)


def buried_count(heavy_xyz, xyz, radius):
    """Heavy atoms of the full receptor within `radius` of each typed atom, the atom itself excluded (the
    buriedness of receptor_annotation, src/adock/fit/terms.py:204-221)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(np.asarray(heavy_xyz, np.float64))
    return np.array([len(n) for n in tree.query_ball_point(np.asarray(xyz, np.float64), radius)],
                    np.float32) - 1


# ------------------------------------------------------------------ basis functions and exact features
def _ramp(x, good, bad):
    """1 at `good`, 0 at `bad`, linear between, clamped (either orientation)."""
    return torch.clamp((bad - x) / (bad - good), 0.0, 1.0)


def basis(r, d, spec):
    """[..., N_BASIS] protein-ligand basis values at distances r and surface distances d, zero beyond the cutoff
    (inter parameters spec.iprm and rep_cap) (src/adock/fit/terms.py:239-254)."""
    p = spec.iprm
    g1 = torch.exp(-((d - p["g1_off"]) / p["g1_w"]) ** 2)
    g2 = torch.exp(-((d - p["g2_off"]) / p["g2_w"]) ** 2)
    rep = torch.clamp(d - p["rep_off"], max=0.0) ** 2
    if spec.rep_cap is not None:
        rep = torch.clamp(rep, max=spec.rep_cap)
    hyd = _ramp(d, p["hyd_good"], p["hyd_bad"])
    hb = _ramp(d, p["hb_good"], p["hb_bad"])
    ion = 1.0 / torch.clamp(r, min=spec.ion_rmin) ** 2 - 1.0 / spec.cutoff ** 2
    b = torch.stack([g1, g2, rep, hyd, hb, ion], -1)
    return b * (r < spec.cutoff).unsqueeze(-1).to(b.dtype)


def inter_features(X, lig_ext, rec_xyz, rec_ext, spec, I=None, chunk=4_000_000):
    """Inter-molecular features [P, N_INTER] of poses X [P, A, 3] (ligand heavy atoms, ext types lig_ext [A]) against
    receptor atoms rec_xyz [N, 3] with ext types rec_ext [N] (src/adock/fit/terms.py:257-280)."""
    dev = X.device
    I = torch.as_tensor(spec.indicator() if I is None else I, device=dev)
    lig_ext = torch.as_tensor(lig_ext, device=dev)
    rec_ext = torch.as_tensor(rec_ext, device=dev)
    rec_xyz = torch.as_tensor(rec_xyz, device=dev, dtype=X.dtype)
    P, A = X.shape[:2]
    F = torch.zeros(P, N_INTER, device=dev, dtype=X.dtype)
    # receptor atoms that can reach any pose within the cutoff
    lo = X.reshape(-1, 3).min(0).values - spec.cutoff
    hi = X.reshape(-1, 3).max(0).values + spec.cutoff
    near = ((rec_xyz >= lo) & (rec_xyz <= hi)).all(-1)
    rx, rt = rec_xyz[near], rec_ext[near]
    if len(rx) == 0:
        return F
    rad_l = torch.as_tensor(spec.xs_radius[spec.lig_xs], device=dev, dtype=X.dtype)[lig_ext]
    rad_r = torch.as_tensor(spec.rec_radius, device=dev, dtype=X.dtype)[rt]
    Ip = I[rt[None, :], lig_ext[:, None]].to(X.dtype)          # [A, N, K, M]
    step = max(1, chunk // max(1, A * len(rx)))
    for p0 in range(0, P, step):
        Xs = X[p0:p0 + step]
        r = torch.linalg.norm(Xs[:, :, None, :] - rx[None, None], dim=-1)
        d = r - rad_l[None, :, None] - rad_r[None, None, :]
        F[p0:p0 + step] = torch.einsum("pank,ankm->pm", basis(r, d, spec), Ip)
    return F


def intra_features(X, lig_xs, pairs, spec):
    """The five Vina terms [P, 5] summed over the ligand's pairs [Q, 2] (XS types lig_xs), always with the base
    parameters (src/adock/fit/terms.py:283-295)."""
    from ..scoring import pair_terms
    dev = X.device
    pairs = np.asarray(pairs, np.int64).reshape(-1, 2)
    if len(pairs) == 0:
        return torch.zeros(X.shape[0], 5, device=dev, dtype=X.dtype)
    pi = torch.as_tensor(pairs, device=dev)
    r = torch.linalg.norm(X[:, pi[:, 0]] - X[:, pi[:, 1]], dim=-1)
    lt = torch.as_tensor(np.asarray(lig_xs, np.int64), device=dev)
    T = lambda a: torch.as_tensor(a, device=dev)
    t = pair_terms(r, lt[pi[:, 0]].expand_as(r), lt[pi[:, 1]].expand_as(r), T(spec.xs_radius), T(spec.xs_hyd),
                   T(spec.xs_don), T(spec.xs_acc), spec.prm_tuple())
    return t.sum(1)


# ------------------------------------------------------------------ the potential object the search reads
class FittedPotential:
    """Stand-in for PairPotential in a Problem (src/adock/fit/terms.py:299-350).

    Ligand types are compact indices 0..TL-1 into lig_ext_list; receptor types in pair calls are TL + receptor ext
    type. energy(r, ti, tj) is the fitted inter energy when tj >= TL and the Vina-form intra energy with the intra
    weights otherwise. `p` carries what the search kernels read: the base parameters, the intra weights as w_g1..w_hb
    and the per-compact-type radius / flag arrays. The search kernels' exact pairwise mode does not implement this
    function; inter_exact (exact rescoring of final poses) runs the fitted-function kernel adock_inter_exact_fit."""

    def __init__(self, spec, weights, lig_ext_list, device="cpu"):
        w = np.asarray(weights, np.float32)
        if w.shape != (len(TERM_NAMES),):
            raise ValueError(f"expected {len(TERM_NAMES)} weights")
        self.spec = spec
        self.lig_ext_list = np.asarray(lig_ext_list, np.int64)
        self.TL = len(self.lig_ext_list)
        xs = spec.lig_xs[self.lig_ext_list]
        T = lambda a: torch.as_tensor(a, device=device)
        self.radius = T(spec.xs_radius[xs])
        self.hyd, self.don, self.acc = T(spec.xs_hyd[xs]), T(spec.xs_don[xs]), T(spec.xs_acc[xs])
        self.w_intra = T(w[N_INTER:])
        self.W = T(spec.pair_table(w[:N_INTER])[:, self.lig_ext_list, :])      # [n_rt, TL, 6]
        self.rad_rec = T(spec.rec_radius)
        wi = [float(v) for v in w[N_INTER:]]
        self.p = {"rot_weight": spec.rot_weight, "name": "fitted", "cutoff": spec.cutoff, **spec.prm,
                  "w_g1": wi[0], "w_g2": wi[1], "w_rep": wi[2], "w_hyd": wi[3], "w_hb": wi[4],
                  "radius": spec.xs_radius[xs].astype(np.float32),
                  "hydrophobic": spec.xs_hyd[xs].astype(np.int32), "donor": spec.xs_don[xs].astype(np.int32),
                  "acceptor": spec.xs_acc[xs].astype(np.int32)}

    def to(self, device):
        for k in ("radius", "hyd", "don", "acc", "w_intra", "W", "rad_rec"):
            setattr(self, k, getattr(self, k).to(device))
        return self

    def prm(self):
        return self.spec.prm_tuple()

    def weights(self, device=None):
        return self.w_intra if device is None else self.w_intra.to(device)

    def inter_exact(self, X, lig_types, rec_idx, rec):
        """Exact pairwise inter energy [L, M] of cuda poses X [L, M, A, 3] by the HIP kernel (gridfit.inter_exact_fit);
        rec = (xyz [R, N, 3], pair types TL + ext type [R, N], mask [R, N]) as Problem.rec."""
        from .gridfit import inter_exact_fit
        rx, rt, rm = (t.to(X.device) for t in rec)
        rext = torch.where(rm, rt - self.TL, torch.full_like(rt, -1))
        return inter_exact_fit(X, lig_types, rec_idx, rx, rext, self)

    def energy(self, r, ti, tj):
        from ..scoring import pair_terms
        inter = tj >= self.TL
        if bool(inter.all()):
            ta = tj - self.TL
            dev = r.device
            d = r - self.radius.to(dev, r.dtype)[ti] - self.rad_rec.to(dev, r.dtype)[ta]
            return (basis(r, d, self.spec) * self.W.to(dev, r.dtype)[ta, ti]).sum(-1)
        if bool(inter.any()):
            raise ValueError("energy(): inter and intra pairs mixed in one call")
        dev = r.device
        t = pair_terms(r, ti, tj, self.radius.to(dev), self.hyd.to(dev), self.don.to(dev), self.acc.to(dev),
                       self.prm())
        return (t * self.w_intra.to(dev, t.dtype)).sum(-1)
