// TEST INFRASTRUCTURE ONLY. CPU restatement of the reference docking hot path, used by tests/ as the parity checker
// and by bench.py as the CPU baseline ("port"). Nothing in the product library links or calls this file.
//
// Every routine is templated on the arithmetic type R: R = double is the FP64 oracle the HIP kernels are checked
// against; R = float is the single-precision restatement timed as the CPU baseline. Inputs are always the float32
// arrays the GPU path receives. Discrete decisions that the reference takes in float32 and that the HIP kernels take in
// float32 as well (grid-point coordinates and the receptor cutoff test of the grid builder) are taken here in float32
// with the same fused-multiply-add nesting, so both sides see the same set of in-cutoff atoms.
//
// Followed reference code (paths relative to the reference repository):
//   pair_e ........................ src/adock/mckernel.py:44-64 (also scoring.py:18-31, grid.py:40-59)
//   build_grid .................... src/adock/grid.py:14-71 (tiling of cpukernel.py:102-137 is not needed on a CPU
//                                    oracle: every point sees every atom, in receptor order)
//   build_grid_fit ................ src/adock/fit/gridfit.py:13-58 (basis functions as fit/terms.py:239-254)
//   inter_exact_fit ............... src/adock/search.py:73-89 with FittedPotential.energy, fit/terms.py:339-344
//   smix / urand / rand_sphere .... src/adock/mckernel.py:71-90
//   quat_inc ...................... src/adock/mckernel.py:93-105
//   eval_eg ....................... src/adock/mckernel.py:108-244
//   conf_inc ...................... src/adock/mckernel.py:256-265
//   bfgs .......................... src/adock/mckernel.py:268-326
//   mc chain ...................... src/adock/mckernel.py:328-408
//   eg ............................ src/adock/mckernel.py:411-447
//   stage_a ....................... src/adock/mckernel.py:451-521
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <omp.h>
#include <vector>

namespace {

struct LigArrays {            // packed ligand layout of src/adock/mckernel.py:547-592
    const float* ref; const int* typ; const unsigned char* tm; const int* ax; const float* axp; const float* axu;
    const int* pr; const int* na; const int* nt; const int* np; const float* rg; int P_MAX; const int* gslot;
    int AM, KM;
};
struct GridArrays {           // GridSet flat layout of src/adock/grid.py:197-247
    const float* grid; const float* origin; const float* lo; const float* hi; const int* rec_idx;
    const long long* goffs; const long long* gdims; float spacing; float pen;
};
struct SFArrays { const float* prm; const float* rad; const int* hyd; const int* don; const int* acc; };
struct RecArrays { const float* rxyz; const int* rtyp; const int* rn; int RN_MAX; };

template <class R> struct SF {
    R g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good, hb_bad, cutoff, w[5];
    const float* rad; const int *hyd, *don, *acc;
    explicit SF(const SFArrays& s) {
        const float* p = s.prm;   // search-kernel order: g1_off..hb_bad, cutoff, w_g1..w_hb
        g1_off = p[0]; g1_w = p[1]; g2_off = p[2]; g2_w = p[3]; rep_off = p[4]; hyd_good = p[5]; hyd_bad = p[6];
        hb_good = p[7]; hb_bad = p[8]; cutoff = p[9];
        for (int i = 0; i < 5; ++i) w[i] = p[10 + i];
        rad = s.rad; hyd = s.hyd; don = s.don; acc = s.acc;
    }
};

inline double r_exp(double x) { return std::exp(x); }
inline float r_exp(float x) { return std::exp(x); }

// Vina-family pair energy and dE/dr
template <class R> R pair_e(R r, int ti, int tj, const SF<R>& s, R& de) {
    R d = r - R(s.rad[ti]) - R(s.rad[tj]);
    R u1 = (d - s.g1_off) / s.g1_w, g1 = r_exp(-u1 * u1);
    R u2 = (d - s.g2_off) / s.g2_w, g2 = r_exp(-u2 * u2);
    R ep = s.w[0] * g1 + s.w[1] * g2;
    de = s.w[0] * (R(-2) * u1 / s.g1_w) * g1 + s.w[1] * (R(-2) * u2 / s.g2_w) * g2;
    R dr = d - s.rep_off;
    if (dr < 0) { ep += s.w[2] * dr * dr; de += s.w[2] * R(2) * dr; }
    if (s.hyd[ti] && s.hyd[tj]) {
        R v = (s.hyd_bad - d) / (s.hyd_bad - s.hyd_good);
        if (v >= 1) ep += s.w[3];
        else if (v > 0) { ep += s.w[3] * v; de += s.w[3] * (R(-1) / (s.hyd_bad - s.hyd_good)); }
    }
    if ((s.don[ti] && s.acc[tj]) || (s.acc[ti] && s.don[tj])) {
        R v = (s.hb_bad - d) / (s.hb_bad - s.hb_good);
        if (v >= 1) ep += s.w[4];
        else if (v > 0) { ep += s.w[4] * v; de += s.w[4] * (R(-1) / (s.hb_bad - s.hb_good)); }
    }
    return ep;
}

// distance (in A) from d to the nearest kink of the slope terms of this type pair; large if none
template <class R> R kink_margin(R d, int ti, int tj, const SF<R>& s) {
    R m = R(1e30);
    if (s.hyd[ti] && s.hyd[tj]) m = std::min(m, std::min(std::fabs(d - s.hyd_good), std::fabs(d - s.hyd_bad)));
    if ((s.don[ti] && s.acc[tj]) || (s.acc[ti] && s.don[tj]))
        m = std::min(m, std::min(std::fabs(d - s.hb_good), std::fabs(d - s.hb_bad)));
    return m;
}

// ---------------------------------------------------------------- grid builder (src/adock/grid.py:14-71)
template <class R>
void build_grid(const float* rxyz, const int* rtype, int n_rec, const float* radius, const int* hyd, const int* don,
                const int* acc, const int* lig_types, int n_lt, float ox, float oy, float oz, float spacing, int nx,
                int ny, int nz, const float* prm, int per_term, R* out) {
    // prm in grid order: cutoff, g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad, hb_good, hb_bad, w_g1..w_hb
    const long long npts = (long long)nx * ny * nz;
    const float cut2f = prm[0] * prm[0];
    const R g1o = prm[1], g1w = prm[2], g2o = prm[3], g2w = prm[4], repo = prm[5];
    const R hg = prm[6], hbad = prm[7], bg = prm[8], bbad = prm[9];
#pragma omp parallel for schedule(dynamic, 64)
    for (long long p = 0; p < npts; ++p) {
        const int k = (int)(p % nz), j = (int)((p / nz) % ny), i = (int)(p / ((long long)ny * nz));
        const float xf = std::fma((float)i, spacing, ox), yf = std::fma((float)j, spacing, oy),
                    zf = std::fma((float)k, spacing, oz);
        R acc_e[5 * 24];
        for (int q = 0; q < 5 * n_lt; ++q) acc_e[q] = 0;
        for (int a = 0; a < n_rec; ++a) {
            const float dxf = rxyz[3 * a] - xf, dyf = rxyz[3 * a + 1] - yf, dzf = rxyz[3 * a + 2] - zf;
            const float r2f = std::fma(dxf, dxf, std::fma(dyf, dyf, dzf * dzf));
            if (!(r2f < cut2f)) continue;
            const R dx = dxf, dy = dyf, dz = dzf;
            const R r = std::sqrt(dx * dx + dy * dy + dz * dz);
            const int ta = rtype[a];
            for (int q = 0; q < n_lt; ++q) {
                const int tl = lig_types[q];
                const R d = r - R(radius[ta]) - R(radius[tl]);
                R u = (d - g1o) / g1w;
                const R g1 = r_exp(-u * u);
                u = (d - g2o) / g2w;
                const R g2 = r_exp(-u * u);
                const R dr = d - repo;
                const R rep = dr < 0 ? dr * dr : R(0);
                R hy = 0, hb = 0;
                if (hyd[ta] && hyd[tl]) hy = std::min(std::max((hbad - d) / (hbad - hg), R(0)), R(1));
                if ((don[ta] && acc[tl]) || (acc[ta] && don[tl])) hb = std::min(std::max((bbad - d) / (bbad - bg), R(0)), R(1));
                acc_e[5 * q + 0] += g1; acc_e[5 * q + 1] += g2; acc_e[5 * q + 2] += rep;
                acc_e[5 * q + 3] += hy; acc_e[5 * q + 4] += hb;
            }
        }
        for (int q = 0; q < n_lt; ++q) {
            if (per_term) {
                for (int m = 0; m < 5; ++m) out[(long long)(m * n_lt + q) * npts + p] = acc_e[5 * q + m];
            } else {
                R e = 0;
                for (int m = 0; m < 5; ++m) e += R(prm[10 + m]) * acc_e[5 * q + m];
                out[(long long)q * npts + p] = e;
            }
        }
    }
}

// ---------------------------------------------------------------- fitted basis (src/adock/fit/terms.py:239-254)
// sum_k W[rt, q, k] * basis_k of one receptor-ligand pair at distance r: gauss1, gauss2, repulsion capped at rep_cap,
// hydrophobic ramp and h-bond ramp of d = r - rad_rec[rt] - rad_lig[q], and the ionic basis
// 1 / max(r, ion_rmin)^2 - 1 / cutoff^2. prm: cutoff, g1_off, g1_w, g2_off, g2_w, rep_off, hyd_good, hyd_bad,
// hb_good, hb_bad, ion_rmin, rep_cap.
template <class R>
struct FitBasis {
    R rc, g1o, g1w, g2o, g2w, repo, hg, hbad, bg, bbad, rmin, cap;
    explicit FitBasis(const float* prm)
        : rc(prm[0]), g1o(prm[1]), g1w(prm[2]), g2o(prm[3]), g2w(prm[4]), repo(prm[5]), hg(prm[6]), hbad(prm[7]),
          bg(prm[8]), bbad(prm[9]), rmin(prm[10]), cap(prm[11]) {}
    R ion(R r) const {
        const R rr = std::max(r, rmin);
        return R(1) / (rr * rr) - R(1) / (rc * rc);
    }
    R pair(const float* w, R d, R ion_r) const {
        R u = (d - g1o) / g1w;
        R sv = R(w[0]) * r_exp(-u * u);
        u = (d - g2o) / g2w;
        sv += R(w[1]) * r_exp(-u * u);
        const R dr = d - repo;
        if (dr < 0) sv += R(w[2]) * std::min(dr * dr, cap);
        sv += R(w[3]) * std::min(std::max((hbad - d) / (hbad - hg), R(0)), R(1));
        sv += R(w[4]) * std::min(std::max((bbad - d) / (bbad - bg), R(0)), R(1));
        sv += R(w[5]) * ion_r;
        return sv;
    }
};

// ---------------------------------------------------------------- fitted grid (src/adock/fit/gridfit.py:13-58)
// out[q, i, j, k] = sum over receptor atoms a (ext type in [0, n_rt)) within the cutoff of the FitBasis pair sum with
// W[rext_a, q].
template <class R>
void build_grid_fit(const float* rxyz, const int* rext, int n_rec, const float* rad_rec, const float* rad_lig,
                    int n_lt, const float* W, int n_rt, float ox, float oy, float oz, float spacing, int nx, int ny,
                    int nz, const float* prm, R* out) {
    const long long npts = (long long)nx * ny * nz;
    const float cut2f = prm[0] * prm[0];
    const FitBasis<R> fb(prm);
#pragma omp parallel for schedule(dynamic, 64)
    for (long long p = 0; p < npts; ++p) {
        const int k = (int)(p % nz), j = (int)((p / nz) % ny), i = (int)(p / ((long long)ny * nz));
        const float xf = std::fma((float)i, spacing, ox), yf = std::fma((float)j, spacing, oy),
                    zf = std::fma((float)k, spacing, oz);
        std::vector<R> e(n_lt, R(0));
        for (int a = 0; a < n_rec; ++a) {
            const int ta = rext[a];
            if (ta < 0 || ta >= n_rt) continue;
            const float dxf = rxyz[3 * a] - xf, dyf = rxyz[3 * a + 1] - yf, dzf = rxyz[3 * a + 2] - zf;
            const float r2f = std::fma(dxf, dxf, std::fma(dyf, dyf, dzf * dzf));
            if (!(r2f < cut2f)) continue;
            const R dx = dxf, dy = dyf, dz = dzf;
            const R r = std::sqrt(dx * dx + dy * dy + dz * dz);
            const R ion = fb.ion(r);
            for (int q = 0; q < n_lt; ++q)
                e[q] += fb.pair(W + ((long long)ta * n_lt + q) * 6, r - R(rad_rec[ta]) - R(rad_lig[q]), ion);
        }
        for (int q = 0; q < n_lt; ++q) out[(long long)q * npts + p] = e[q];
    }
}

// ---------------------------------------------------------------- exact fitted inter energy (src/adock/search.py:73-89)
// out[l, m] = sum over ligand atoms a of pose (l, m) with lig_type in [0, n_lt) and receptor atoms b of set
// rec_idx[l] with rext in [0, n_rt) and |x_b - x_a|^2 < cutoff^2 (decided in float32 with the fused-multiply-add
// nesting of the HIP kernel) of the FitBasis pair sum with W[rext_b, lig_type_a] (FittedPotential.energy,
// src/adock/fit/terms.py:339-344). A complex whose rec_idx lies outside [0, n_sets) gets NaN.
template <class R>
void inter_exact_fit(const float* X, const int* lig_type, int L, int M, int A, const int* rec_idx, const float* rxyz,
                     const int* rext, int n_sets, int NR, const float* rad_rec, const float* rad_lig, int n_lt,
                     const float* W, int n_rt, const float* prm, R* out) {
    const float cut2f = prm[0] * prm[0];
    const FitBasis<R> fb(prm);
#pragma omp parallel for schedule(dynamic, 4)
    for (long long lm = 0; lm < (long long)L * M; ++lm) {
        const int l = (int)(lm / M);
        const int ri = rec_idx[l];
        if (ri < 0 || ri >= n_sets) {
            out[lm] = std::nan("");
            continue;
        }
        const float* rx = rxyz + (long long)ri * NR * 3;
        const int* re = rext + (long long)ri * NR;
        R e = 0;
        for (int a = 0; a < A; ++a) {
            const int q = lig_type[(long long)l * A + a];
            if (q < 0 || q >= n_lt) continue;
            const float* xa = X + (lm * A + a) * 3;
            for (int b = 0; b < NR; ++b) {
                const int tb = re[b];
                if (tb < 0 || tb >= n_rt) continue;
                const float dxf = rx[3 * b] - xa[0], dyf = rx[3 * b + 1] - xa[1], dzf = rx[3 * b + 2] - xa[2];
                const float r2f = std::fma(dxf, dxf, std::fma(dyf, dyf, dzf * dzf));
                if (!(r2f < cut2f)) continue;
                const R dx = dxf, dy = dyf, dz = dzf;
                const R r = std::sqrt(dx * dx + dy * dy + dz * dz);
                e += fb.pair(W + ((long long)tb * n_lt + q) * 6, r - R(rad_rec[tb]) - R(rad_lig[q]), fb.ion(r));
            }
        }
        out[lm] = e;
    }
}

// ---------------------------------------------------------------- counter-based RNG (src/adock/mckernel.py:71-90)
inline uint64_t smix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
inline float urand(uint64_t seed, int chain, int step, int slot) {
    uint64_t h = smix(seed ^ smix(((uint64_t)(uint32_t)chain << 32) ^ ((uint64_t)(int64_t)step << 8) ^
                                  (uint64_t)(int64_t)slot));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}
inline void rand_sphere(uint64_t seed, int chain, int step, int slot, float* v) {
    for (int k = 0; k < 32; ++k) {
        float a = 2.f * urand(seed, chain, step, slot + 3 * k) - 1.f;
        float b = 2.f * urand(seed, chain, step, slot + 3 * k + 1) - 1.f;
        float c = 2.f * urand(seed, chain, step, slot + 3 * k + 2) - 1.f;
        if (a * a + b * b + c * c < 1.f) { v[0] = a; v[1] = b; v[2] = c; return; }
    }
    v[0] = v[1] = v[2] = 0.f;
}

// q <- quat(rotvec) * q, normalised
template <class R> void quat_inc(R* q, R rx, R ry, R rz) {
    R ang = std::sqrt(rx * rx + ry * ry + rz * rz);
    R hw, k;
    if (ang > R(1e-6)) { hw = std::cos(R(0.5) * ang); k = std::sin(R(0.5) * ang) / ang; } else { hw = 1; k = R(0.5); }
    R aw = hw, ax = rx * k, ay = ry * k, az = rz * k;
    R bw = q[0], bx = q[1], by = q[2], bz = q[3];
    R w = aw * bw - ax * bx - ay * by - az * bz;
    R x = aw * bx + ax * bw + ay * bz - az * by;
    R y = aw * by - ax * bz + ay * bw + az * bx;
    R z = aw * bz + ax * by - ay * bx + az * bw;
    R nn = R(1) / std::sqrt(w * w + x * x + y * y + z * z);
    q[0] = w * nn; q[1] = x * nn; q[2] = y * nn; q[3] = z * nn;
}

// per-chain context: ligand l, receptor ri
template <class R> struct Ctx {
    const float *ref, *axp, *axu; const int *typ, *ax, *pr, *gslot; const unsigned char* tm;
    int na, nt, np, AM; R rg;
    const float* grid; R o[3], lo[3], hi[3]; int nx, ny, nz; R inv_s, pen;
    const float* rx; const int* rt; int rn; int exact; const int* nb; int nbn;
};

template <class R> Ctx<R> make_ctx(const LigArrays& L, const GridArrays& G, const RecArrays& Rc, int l, int exact) {
    Ctx<R> c;
    const int AM = L.AM, KM = L.KM;
    c.AM = AM;
    c.ref = L.ref + (long long)l * AM * 3; c.typ = L.typ + (long long)l * AM; c.tm = L.tm + (long long)l * KM * AM;
    c.ax = L.ax + (long long)l * KM * 2; c.axp = L.axp + (long long)l * KM * 3; c.axu = L.axu + (long long)l * KM * 3;
    c.pr = L.pr + (long long)l * L.P_MAX * 2; c.gslot = L.gslot + (long long)l * AM;
    c.na = L.na[l]; c.nt = L.nt[l]; c.np = L.np[l]; c.rg = L.rg[l];
    const int ri = G.rec_idx[l];
    c.grid = G.grid + G.goffs[ri];
    c.nx = (int)G.gdims[3 * ri]; c.ny = (int)G.gdims[3 * ri + 1]; c.nz = (int)G.gdims[3 * ri + 2];
    for (int d = 0; d < 3; ++d) { c.o[d] = G.origin[3 * ri + d]; c.lo[d] = G.lo[3 * ri + d]; c.hi[d] = G.hi[3 * ri + d]; }
    c.inv_s = R(1) / R(G.spacing); c.pen = G.pen;
    c.rx = Rc.rxyz ? Rc.rxyz + (long long)ri * Rc.RN_MAX * 3 : nullptr;
    c.rt = Rc.rtyp ? Rc.rtyp + (long long)ri * Rc.RN_MAX : nullptr;
    c.rn = Rc.rn ? Rc.rn[ri] : 0;
    c.exact = exact; c.nb = nullptr; c.nbn = 0;
    return c;
}

constexpr int AMX = 256, KMX = 64;

// energy + gradient of state s (7 + nt) -> g (6 + nt). X (optional) receives world coordinates [na, 3].
// margin (optional) receives the smallest distance of the pose to a discontinuity of E or of dE: a cell face or the
// clamp limits of the trilinear lookup (A), the pair cutoff (A), a slope-term kink (A).
template <class R>
R eval_eg(const R* s, R* g, const Ctx<R>& c, const SF<R>& sf, R* Xout = nullptr, R* margin = nullptr) {
    R X[AMX * 3], F[AMX * 3], pw[KMX * 3], uw[KMX * 3];
    R qw = s[3], qx = s[4], qy = s[5], qz = s[6];
    R nq = R(1) / std::sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
    qw *= nq; qx *= nq; qy *= nq; qz *= nq;
    const R R00 = 1 - 2 * (qy * qy + qz * qz), R01 = 2 * (qx * qy - qz * qw), R02 = 2 * (qx * qz + qy * qw);
    const R R10 = 2 * (qx * qy + qz * qw), R11 = 1 - 2 * (qx * qx + qz * qz), R12 = 2 * (qy * qz - qx * qw);
    const R R20 = 2 * (qx * qz - qy * qw), R21 = 2 * (qy * qz + qx * qw), R22 = 1 - 2 * (qx * qx + qy * qy);
    const long long n3 = (long long)c.nx * c.ny * c.nz;
    const int ny = c.ny, nz = c.nz;
    const R fmx[3] = {R(c.nx - 1) - R(1e-4), R(c.ny - 1) - R(1e-4), R(c.nz - 1) - R(1e-4)};
    const R sp = R(1) / c.inv_s;
    R mg = R(1e30);
    R e = 0;
    const R cut = sf.cutoff;
    for (int a = 0; a < c.na; ++a) {
        R px = c.ref[3 * a], py = c.ref[3 * a + 1], pz = c.ref[3 * a + 2];
        for (int k = 0; k < c.nt; ++k) {
            if (c.tm[k * c.AM + a]) {
                R th = s[7 + k];
                R cs = std::cos(th), sn = std::sin(th);
                R ux = c.axu[3 * k], uy = c.axu[3 * k + 1], uz = c.axu[3 * k + 2];
                R vx = px - R(c.axp[3 * k]), vy = py - R(c.axp[3 * k + 1]), vz = pz - R(c.axp[3 * k + 2]);
                R d = ux * vx + uy * vy + uz * vz;
                R cx = uy * vz - uz * vy, cy = uz * vx - ux * vz, cz = ux * vy - uy * vx;
                px = R(c.axp[3 * k]) + vx * cs + cx * sn + ux * d * (1 - cs);
                py = R(c.axp[3 * k + 1]) + vy * cs + cy * sn + uy * d * (1 - cs);
                pz = R(c.axp[3 * k + 2]) + vz * cs + cz * sn + uz * d * (1 - cs);
            }
        }
        const R X0 = R00 * px + R01 * py + R02 * pz + s[0];
        const R X1 = R10 * px + R11 * py + R12 * pz + s[1];
        const R X2 = R20 * px + R21 * py + R22 * pz + s[2];
        X[3 * a] = X0; X[3 * a + 1] = X1; X[3 * a + 2] = X2;
        const R xx[3] = {X0, X1, X2};
        R fw[3], inside[3]; long long i0[3];
        for (int d = 0; d < 3; ++d) {
            R f = (xx[d] - c.o[d]) * c.inv_s;
            R fc = std::min(std::max(f, R(0)), fmx[d]);
            inside[d] = (f > 0 && f < fmx[d]) ? R(1) : R(0);
            long long ii = (long long)std::floor(fc);
            i0[d] = ii; fw[d] = fc - R(ii);
            if (margin && !c.exact) {
                R frac = f - std::floor(f);
                mg = std::min(mg, std::min(frac, R(1) - frac) * sp);
                mg = std::min(mg, std::min(std::fabs(f), std::fabs(f - fmx[d])) * sp);
            }
        }
        if (c.exact) {
            R ea = 0, gx = 0, gy = 0, gz = 0;
            const int ta = c.typ[a];
            const R cut2 = cut * cut;
            const int nloop = c.nb ? c.nbn : c.rn;
            for (int bb = 0; bb < nloop; ++bb) {
                const int b = c.nb ? c.nb[bb] : bb;
                R rx = X0 - R(c.rx[3 * b]), ry = X1 - R(c.rx[3 * b + 1]), rz = X2 - R(c.rx[3 * b + 2]);
                R r2 = rx * rx + ry * ry + rz * rz;
                if (margin) {
                    R rr = std::sqrt(r2);
                    mg = std::min(mg, std::fabs(rr - cut));
                    if (rr < cut) mg = std::min(mg, kink_margin(rr - R(sf.rad[ta]) - R(sf.rad[c.rt[b]]), ta, c.rt[b], sf));
                }
                if (r2 >= cut2) continue;
                R r = std::sqrt(r2), de;
                ea += pair_e(r, ta, c.rt[b], sf, de);
                R sc = de / std::max(r, R(1e-6));
                gx += sc * rx; gy += sc * ry; gz += sc * rz;
            }
            R gg[3] = {gx, gy, gz};
            for (int d = 0; d < 3; ++d) {
                R lo = std::max(c.lo[d] - xx[d], R(0)), hi = std::max(xx[d] - c.hi[d], R(0));
                ea += c.pen * (lo * lo + hi * hi);
                gg[d] += c.pen * (R(-2) * lo + R(2) * hi);
            }
            e += ea;
            F[3 * a] = gg[0]; F[3 * a + 1] = gg[1]; F[3 * a + 2] = gg[2];
            continue;
        }
        const float* G = c.grid + (long long)c.gslot[a] * n3;
        const long long b000 = (i0[0] * ny + i0[1]) * nz + i0[2];
        const R v000 = G[b000], v001 = G[b000 + 1], v010 = G[b000 + nz], v011 = G[b000 + nz + 1];
        const long long bx = b000 + (long long)ny * nz;
        const R v100 = G[bx], v101 = G[bx + 1], v110 = G[bx + nz], v111 = G[bx + nz + 1];
        const R wx = fw[0], wy = fw[1], wz = fw[2];
        const R c00 = v000 * (1 - wz) + v001 * wz, c01 = v010 * (1 - wz) + v011 * wz;
        const R c10 = v100 * (1 - wz) + v101 * wz, c11 = v110 * (1 - wz) + v111 * wz;
        const R c0 = c00 * (1 - wy) + c01 * wy, c1 = c10 * (1 - wy) + c11 * wy;
        R ea = c0 * (1 - wx) + c1 * wx;
        const R dx = (c1 - c0) * c.inv_s * inside[0];
        const R dy = ((c01 - c00) * (1 - wx) + (c11 - c10) * wx) * c.inv_s * inside[1];
        const R dz = (((v001 - v000) * (1 - wy) + (v011 - v010) * wy) * (1 - wx) +
                      ((v101 - v100) * (1 - wy) + (v111 - v110) * wy) * wx) * c.inv_s * inside[2];
        R gg[3] = {dx, dy, dz};
        for (int d = 0; d < 3; ++d) {
            R lo = std::max(c.lo[d] - xx[d], R(0)), hi = std::max(xx[d] - c.hi[d], R(0));
            ea += c.pen * (lo * lo + hi * hi);
            gg[d] += c.pen * (R(-2) * lo + R(2) * hi);
        }
        e += ea;
        F[3 * a] = gg[0]; F[3 * a + 1] = gg[1]; F[3 * a + 2] = gg[2];
    }
    for (int q = 0; q < c.np; ++q) {
        const int i = c.pr[2 * q], j = c.pr[2 * q + 1];
        if (i == j) continue;   // padding entry of a scheduled pair list
        R rx = X[3 * i] - X[3 * j], ry = X[3 * i + 1] - X[3 * j + 1], rz = X[3 * i + 2] - X[3 * j + 2];
        R r = std::sqrt(rx * rx + ry * ry + rz * rz);
        if (margin) {
            mg = std::min(mg, std::fabs(r - cut));
            if (r < cut) mg = std::min(mg, kink_margin(r - R(sf.rad[c.typ[i]]) - R(sf.rad[c.typ[j]]), c.typ[i], c.typ[j], sf));
        }
        if (r >= cut) continue;
        R de;
        e += pair_e(r, c.typ[i], c.typ[j], sf, de);
        R sc = de / std::max(r, R(1e-6));
        F[3 * i] += sc * rx; F[3 * i + 1] += sc * ry; F[3 * i + 2] += sc * rz;
        F[3 * j] -= sc * rx; F[3 * j + 1] -= sc * ry; F[3 * j + 2] -= sc * rz;
    }
    for (int k = 0; k < c.nt; ++k) {
        const int a = c.ax[2 * k], b = c.ax[2 * k + 1];
        R ux = X[3 * b] - X[3 * a], uy = X[3 * b + 1] - X[3 * a + 1], uz = X[3 * b + 2] - X[3 * a + 2];
        R nn = R(1) / std::sqrt(std::max(ux * ux + uy * uy + uz * uz, R(1e-12)));
        uw[3 * k] = ux * nn; uw[3 * k + 1] = uy * nn; uw[3 * k + 2] = uz * nn;
        pw[3 * k] = X[3 * a]; pw[3 * k + 1] = X[3 * a + 1]; pw[3 * k + 2] = X[3 * a + 2];
        g[6 + k] = 0;
    }
    R gt0 = 0, gt1 = 0, gt2 = 0, tq0 = 0, tq1 = 0, tq2 = 0;
    for (int a = 0; a < c.na; ++a) {
        const R fx = F[3 * a], fy = F[3 * a + 1], fz = F[3 * a + 2];
        const R X0 = X[3 * a], X1 = X[3 * a + 1], X2 = X[3 * a + 2];
        gt0 += fx; gt1 += fy; gt2 += fz;
        const R dx = X0 - s[0], dy = X1 - s[1], dz = X2 - s[2];
        tq0 += dy * fz - dz * fy; tq1 += dz * fx - dx * fz; tq2 += dx * fy - dy * fx;
        for (int k = 0; k < c.nt; ++k) {
            if (c.tm[k * c.AM + a]) {
                const R vx = X0 - pw[3 * k], vy = X1 - pw[3 * k + 1], vz = X2 - pw[3 * k + 2];
                const R cx = vy * fz - vz * fy, cy = vz * fx - vx * fz, cz = vx * fy - vy * fx;
                g[6 + k] += cx * uw[3 * k] + cy * uw[3 * k + 1] + cz * uw[3 * k + 2];
            }
        }
    }
    g[0] = gt0; g[1] = gt1; g[2] = gt2; g[3] = tq0; g[4] = tq1; g[5] = tq2;
    if (Xout) std::memcpy(Xout, X, sizeof(R) * 3 * c.na);
    if (margin) *margin = mg;
    return e;
}

template <class R> R dot(const R* a, const R* b, int D) { R s = 0; for (int i = 0; i < D; ++i) s += a[i] * b[i]; return s; }

// xn = x (+) alpha p
template <class R> void conf_inc(R* xn, const R* x, const R* p, R alpha, int nt) {
    xn[0] = x[0] + alpha * p[0]; xn[1] = x[1] + alpha * p[1]; xn[2] = x[2] + alpha * p[2];
    xn[3] = x[3]; xn[4] = x[4]; xn[5] = x[5]; xn[6] = x[6];
    quat_inc(xn + 3, alpha * p[3], alpha * p[4], alpha * p[5]);
    for (int k = 0; k < nt; ++k) xn[7 + k] = x[7 + k] + alpha * p[6 + k];
}

template <class R> struct Counters { long long evals = 0; };

// Vina-style BFGS on state x (in/out); returns final energy
template <class R> R bfgs(R* x, int iters, const Ctx<R>& c, const SF<R>& sf, long long* evals) {
    const R EPS = R(1.1920929e-07f);
    const int D = 6 + c.nt, S = 7 + c.nt;
    R H[(6 + KMX) * (6 + KMX)], g[6 + KMX], gn[6 + KMX], go[6 + KMX], p[6 + KMX], y[6 + KMX], mhy[6 + KMX];
    R xn[7 + KMX], xo[7 + KMX];
    R f0 = eval_eg(x, g, c, sf); ++*evals;
    const R f_orig = f0;
    std::memcpy(xo, x, sizeof(R) * S); std::memcpy(go, g, sizeof(R) * D);
    for (int i = 0; i < D * D; ++i) H[i] = (i / D == i % D) ? R(1) : R(0);
    for (int step = 0; step < iters; ++step) {
        for (int i = 0; i < D; ++i) { R s = 0; for (int j = 0; j < D; ++j) s += H[i * D + j] * g[j]; p[i] = -s; }
        const R pg = dot(p, g, D);
        R alpha = 1, f1 = 0;
        for (int trial = 0; trial < 10; ++trial) {
            conf_inc(xn, x, p, alpha, c.nt);
            f1 = eval_eg(xn, gn, c, sf); ++*evals;
            if (f1 - f0 < R(1e-4) * alpha * pg) break;
            alpha *= R(0.5);
        }
        for (int i = 0; i < D; ++i) y[i] = gn[i] - g[i];
        f0 = f1;
        std::memcpy(x, xn, sizeof(R) * S);
        const R gg = dot(gn, gn, D);
        std::memcpy(g, gn, sizeof(R) * D);
        if (!(std::sqrt(gg) >= R(1e-5))) break;
        const R yp = dot(y, p, D);
        if (step == 0) {
            const R yy = dot(y, y, D);
            if (std::fabs(yy) > EPS) {
                const R dg = alpha * yp / yy;
                for (int i = 0; i < D * D; ++i) H[i] = (i / D == i % D) ? dg : R(0);
            }
        }
        if (alpha * yp < EPS) continue;
        for (int i = 0; i < D; ++i) { R s = 0; for (int j = 0; j < D; ++j) s += H[i * D + j] * y[j]; mhy[i] = -s; }
        const R yhy = -dot(y, mhy, D);
        const R r = R(1) / (alpha * yp);
        const R c2 = alpha * alpha * (r * r * yhy + r);
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j)
                H[i * D + j] += alpha * r * (mhy[i] * p[j] + mhy[j] * p[i]) + c2 * p[i] * p[j];
    }
    if (!(f0 <= f_orig)) { f0 = f_orig; std::memcpy(x, xo, sizeof(R) * S); }
    return f0;
}

// one MC chain (src/adock/mckernel.py:365-407); cur/best are SD floats, in/out
template <class R>
void mc_chain(const Ctx<R>& c, const SF<R>& sf, int cid, float* gcur, float* gbest, float* cur_e, float* best_e,
              int init_opt, int mc_steps, int bfgs_iters, float temp, float amp_pos, float amp_rot, uint64_t seed,
              int step_offset, long long* evals) {
    const int S = 7 + c.nt;
    const int iters = bfgs_iters > 0 ? bfgs_iters : (25 + c.na) / 3;
    R cur[7 + KMX], best[7 + KMX], x[7 + KMX];
    for (int i = 0; i < S; ++i) { cur[i] = gcur[i]; best[i] = gbest[i]; }
    R ec, eb = *best_e;
    if (init_opt) ec = bfgs(cur, iters, c, sf, evals); else ec = *cur_e;
    if (ec < eb) { eb = ec; std::memcpy(best, cur, sizeof(R) * S); }
    for (int st = 0; st < mc_steps; ++st) {
        const int sg = step_offset + st;
        std::memcpy(x, cur, sizeof(R) * S);
        const int nm = 2 + c.nt;
        const int which = std::min((int)(urand(seed, cid, sg, 0) * (float)nm), nm - 1);
        float v[3];
        if (which == 0) {
            rand_sphere(seed, cid, sg, 10, v);
            x[0] += R(amp_pos) * R(v[0]); x[1] += R(amp_pos) * R(v[1]); x[2] += R(amp_pos) * R(v[2]);
        } else if (which == 1) {
            rand_sphere(seed, cid, sg, 10, v);
            const float scf = amp_rot / std::max((float)c.rg, 1.0f);
            quat_inc(x + 3, R(scf * v[0]), R(scf * v[1]), R(scf * v[2]));
        } else {
            x[7 + which - 2] = R((2.f * urand(seed, cid, sg, 1) - 1.f) * 3.14159265f);
        }
        const R en = bfgs(x, iters, c, sf, evals);
        const bool accept = (en < ec) || (urand(seed, cid, sg, 2) < std::exp(float((ec - en) / R(temp))));
        if (accept) {
            ec = en; std::memcpy(cur, x, sizeof(R) * S);
            if (en < eb) { eb = en; std::memcpy(best, x, sizeof(R) * S); }
        }
    }
    for (int i = 0; i < S; ++i) { gcur[i] = (float)cur[i]; gbest[i] = (float)best[i]; }
    *cur_e = (float)ec; *best_e = (float)eb;
}

template <class R>
void stage_a(const float* conf, const int* gslot, const int* na, int C, const float* quats, int Q, const float* trans,
             const int* ntrans, int NT_MAX, const GridArrays& G, int AM, float cap, long long n_blocks, R* best_e,
             int* best_i) {
#pragma omp parallel for schedule(dynamic, 1)
    for (long long b = 0; b < n_blocks; ++b) {
        const int q = (int)(b % Q), c = (int)((b / Q) % C);
        const int l = (int)(b / ((long long)Q * C));
        const int n_at = na[l];
        R qw = quats[4 * q], qx = quats[4 * q + 1], qy = quats[4 * q + 2], qz = quats[4 * q + 3];
        R nq = R(1) / std::sqrt(qw * qw + qx * qx + qy * qy + qz * qz);
        qw *= nq; qx *= nq; qy *= nq; qz *= nq;
        R X[AMX * 3];
        int Sl[AMX];
        for (int a = 0; a < n_at; ++a) {
            const float* p = conf + (((long long)l * C + c) * AM + a) * 3;
            R px = p[0], py = p[1], pz = p[2];
            X[3 * a] = (1 - 2 * (qy * qy + qz * qz)) * px + 2 * (qx * qy - qz * qw) * py + 2 * (qx * qz + qy * qw) * pz;
            X[3 * a + 1] = 2 * (qx * qy + qz * qw) * px + (1 - 2 * (qx * qx + qz * qz)) * py + 2 * (qy * qz - qx * qw) * pz;
            X[3 * a + 2] = 2 * (qx * qz - qy * qw) * px + 2 * (qy * qz + qx * qw) * py + (1 - 2 * (qx * qx + qy * qy)) * pz;
            Sl[a] = gslot[(long long)l * AM + a];
        }
        const int ri = G.rec_idx[l];
        const float* Gp = G.grid + G.goffs[ri];
        const int nx = (int)G.gdims[3 * ri], ny = (int)G.gdims[3 * ri + 1], nz = (int)G.gdims[3 * ri + 2];
        const long long n3 = (long long)nx * ny * nz;
        const R ox = G.origin[3 * ri], oy = G.origin[3 * ri + 1], oz = G.origin[3 * ri + 2];
        const R lx = G.lo[3 * ri], ly = G.lo[3 * ri + 1], lz = G.lo[3 * ri + 2];
        const R hx = G.hi[3 * ri], hy = G.hi[3 * ri + 1], hz = G.hi[3 * ri + 2];
        const R inv = R(1) / R(G.spacing), pen = G.pen;
        R be = R(3.4e38); int bi = 0;
        for (int ti = 0; ti < ntrans[l]; ++ti) {
            const float* T = trans + ((long long)l * NT_MAX + ti) * 3;
            R e = 0;
            for (int a = 0; a < n_at; ++a) {
                const R x = X[3 * a] + R(T[0]), y = X[3 * a + 1] + R(T[1]), z = X[3 * a + 2] + R(T[2]);
                const int i = std::min(std::max((int)std::nearbyint((x - ox) * inv), 0), nx - 1);
                const int j = std::min(std::max((int)std::nearbyint((y - oy) * inv), 0), ny - 1);
                const int k = std::min(std::max((int)std::nearbyint((z - oz) * inv), 0), nz - 1);
                R ea = Gp[(long long)Sl[a] * n3 + ((long long)i * ny + j) * nz + k];
                const R dx0 = std::max(lx - x, R(0)), dx1 = std::max(x - hx, R(0));
                const R dy0 = std::max(ly - y, R(0)), dy1 = std::max(y - hy, R(0));
                const R dz0 = std::max(lz - z, R(0)), dz1 = std::max(z - hz, R(0));
                ea += pen * (dx0 * dx0 + dx1 * dx1 + dy0 * dy0 + dy1 * dy1 + dz0 * dz0 + dz1 * dz1);
                e += std::min(ea, R(cap));
            }
            if (e < be) { be = e; bi = ti; }
        }
        best_e[b] = be; best_i[b] = bi;
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------- C ABI (ctypes)
#define LIG_PARAMS                                                                                                  \
    const float *ref, const int *typ, const unsigned char *tm, const int *ax, const float *axp, const float *axu,   \
        const int *pr, const int *na, const int *nt, const int *np, const float *rg, int P_MAX, const int *gslot,   \
        int AM, int KM
#define GRID_PARAMS                                                                                                 \
    const float *grid, const float *origin, const float *lo, const float *hi, const int *rec_idx,                   \
        const long long *goffs, const long long *gdims, float spacing, float pen
#define SF_PARAMS const float *prm, const float *rad, const int *hyd, const int *don, const int *acc
#define REC_PARAMS const float *rxyz, const int *rtyp, const int *rn, int RN_MAX
#define PACK_ALL                                                                                                    \
    LigArrays L{ref, typ, tm, ax, axp, axu, pr, na, nt, np, rg, P_MAX, gslot, AM, KM};                              \
    GridArrays G{grid, origin, lo, hi, rec_idx, goffs, gdims, spacing, pen};                                         \
    SFArrays SA{prm, rad, hyd, don, acc};                                                                            \
    RecArrays RC{rxyz, rtyp, rn, RN_MAX};

extern "C" {

void oracle_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int oracle_get_threads(void) { return omp_get_max_threads(); }
int oracle_limits(int* amx, int* kmx) { *amx = AMX; *kmx = KMX; return 0; }

unsigned long long oracle_smix(unsigned long long x) { return smix(x); }
float oracle_urand(unsigned long long seed, int chain, int step, int slot) { return urand(seed, chain, step, slot); }

void oracle_build_grid_f64(const float* rxyz, const int* rtype, int n_rec, const float* radius, const int* hyd,
                           const int* don, const int* acc, const int* lig_types, int n_lt, float ox, float oy,
                           float oz, float spacing, int nx, int ny, int nz, const float* prm, int per_term,
                           double* out) {
    build_grid<double>(rxyz, rtype, n_rec, radius, hyd, don, acc, lig_types, n_lt, ox, oy, oz, spacing, nx, ny, nz,
                       prm, per_term, out);
}
void oracle_build_grid_f32(const float* rxyz, const int* rtype, int n_rec, const float* radius, const int* hyd,
                           const int* don, const int* acc, const int* lig_types, int n_lt, float ox, float oy,
                           float oz, float spacing, int nx, int ny, int nz, const float* prm, int per_term,
                           float* out) {
    build_grid<float>(rxyz, rtype, n_rec, radius, hyd, don, acc, lig_types, n_lt, ox, oy, oz, spacing, nx, ny, nz,
                      prm, per_term, out);
}

void oracle_build_grid_fit_f64(const float* rxyz, const int* rext, int n_rec, const float* rad_rec,
                               const float* rad_lig, int n_lt, const float* W, int n_rt, float ox, float oy, float oz,
                               float spacing, int nx, int ny, int nz, const float* prm, double* out) {
    build_grid_fit<double>(rxyz, rext, n_rec, rad_rec, rad_lig, n_lt, W, n_rt, ox, oy, oz, spacing, nx, ny, nz, prm,
                           out);
}
void oracle_build_grid_fit_f32(const float* rxyz, const int* rext, int n_rec, const float* rad_rec,
                               const float* rad_lig, int n_lt, const float* W, int n_rt, float ox, float oy, float oz,
                               float spacing, int nx, int ny, int nz, const float* prm, float* out) {
    build_grid_fit<float>(rxyz, rext, n_rec, rad_rec, rad_lig, n_lt, W, n_rt, ox, oy, oz, spacing, nx, ny, nz, prm,
                          out);
}

void oracle_inter_exact_fit_f64(const float* X, const int* lig_type, int L, int M, int A, const int* rec_idx,
                                const float* rxyz, const int* rext, int n_sets, int NR, const float* rad_rec,
                                const float* rad_lig, int n_lt, const float* W, int n_rt, const float* prm,
                                double* out) {
    inter_exact_fit<double>(X, lig_type, L, M, A, rec_idx, rxyz, rext, n_sets, NR, rad_rec, rad_lig, n_lt, W, n_rt,
                            prm, out);
}

// energy + gradient for N states; nbr/nbr_n optional (exact mode). Outputs FP64 e [N], g [N, 6+KM], X [N, AM, 3],
// margin [N] (any may be NULL).
void oracle_eg_f64(LIG_PARAMS, GRID_PARAMS, SF_PARAMS, REC_PARAMS, int exact, const int* chain_lig, int n,
                   const float* state, const int* nbr, const int* nbr_n, int NB_MAX, double* e_out, double* g_out,
                   double* x_out, double* margin) {
    PACK_ALL
    const SF<double> sf(SA);
    const int SD = 7 + KM, DM = 6 + KM;
#pragma omp parallel for schedule(dynamic, 4)
    for (int ch = 0; ch < n; ++ch) {
        Ctx<double> c = make_ctx<double>(L, G, RC, chain_lig[ch], exact);
        if (nbr) { c.nb = nbr + (long long)ch * NB_MAX; c.nbn = nbr_n[ch]; }
        double s[7 + KMX], g[6 + KMX], X[AMX * 3], mg;
        for (int i = 0; i < SD; ++i) s[i] = state[(long long)ch * SD + i];
        const double e = eval_eg(s, g, c, sf, X, &mg);
        if (e_out) e_out[ch] = e;
        if (g_out) for (int i = 0; i < 6 + c.nt; ++i) g_out[(long long)ch * DM + i] = g[i];
        if (x_out) for (int i = 0; i < 3 * c.na; ++i) x_out[(long long)ch * AM * 3 + i] = X[i];
        if (margin) margin[ch] = mg;
    }
}

void oracle_eg_f32(LIG_PARAMS, GRID_PARAMS, SF_PARAMS, REC_PARAMS, int exact, const int* chain_lig, int n,
                   const float* state, float* e_out, float* g_out) {
    PACK_ALL
    const SF<float> sf(SA);
    const int SD = 7 + KM, DM = 6 + KM;
#pragma omp parallel for schedule(dynamic, 4)
    for (int ch = 0; ch < n; ++ch) {
        Ctx<float> c = make_ctx<float>(L, G, RC, chain_lig[ch], exact);
        float s[7 + KMX], g[6 + KMX];
        for (int i = 0; i < SD; ++i) s[i] = state[(long long)ch * SD + i];
        const float e = eval_eg(s, g, c, sf);
        e_out[ch] = e;
        for (int i = 0; i < 6 + c.nt; ++i) g_out[(long long)ch * DM + i] = g[i];
    }
}

// MC chains, in place on cur/cur_e/best/best_e (float32 like the reference buffers); evals_out (optional) [N] counts
// eval_eg calls per chain. use_f64 selects the FP64 oracle or the FP32 restatement.
void oracle_mc(int use_f64, LIG_PARAMS, GRID_PARAMS, SF_PARAMS, REC_PARAMS, int exact, const int* chain_lig,
               const int* chain_id, int n_chains, const int* nbr, const int* nbr_n, int NB_MAX, float* cur,
               float* cur_e, float* best, float* best_e, int init_opt, int mc_steps, int bfgs_iters, float temp,
               float amp_pos, float amp_rot, unsigned long long seed, int step_offset, long long* evals_out) {
    PACK_ALL
    const int SD = 7 + KM;
#pragma omp parallel for schedule(dynamic, 1)
    for (int ch = 0; ch < n_chains; ++ch) {
        long long ev = 0;
        if (use_f64) {
            const SF<double> sf(SA);
            Ctx<double> c = make_ctx<double>(L, G, RC, chain_lig[ch], exact);
            if (nbr) { c.nb = nbr + (long long)ch * NB_MAX; c.nbn = nbr_n[ch]; }
            mc_chain(c, sf, chain_id[ch], cur + (long long)ch * SD, best + (long long)ch * SD, cur_e + ch, best_e + ch,
                     init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot, seed, step_offset, &ev);
        } else {
            const SF<float> sf(SA);
            Ctx<float> c = make_ctx<float>(L, G, RC, chain_lig[ch], exact);
            if (nbr) { c.nb = nbr + (long long)ch * NB_MAX; c.nbn = nbr_n[ch]; }
            mc_chain(c, sf, chain_id[ch], cur + (long long)ch * SD, best + (long long)ch * SD, cur_e + ch, best_e + ch,
                     init_opt, mc_steps, bfgs_iters, temp, amp_pos, amp_rot, seed, step_offset, &ev);
        }
        if (evals_out) evals_out[ch] = ev;
    }
}

void oracle_stage_a_f64(const float* conf, const int* gslot, const int* na, int C, const float* quats, int Q,
                        const float* trans, const int* ntrans, int NT_MAX, GRID_PARAMS, int AM, float cap,
                        long long n_blocks, double* best_e, int* best_i) {
    GridArrays G{grid, origin, lo, hi, rec_idx, goffs, gdims, spacing, pen};
    stage_a<double>(conf, gslot, na, C, quats, Q, trans, ntrans, NT_MAX, G, AM, cap, n_blocks, best_e, best_i);
}

}  // extern "C"
