// asph_adjoint_ref.cpp - FP64 host adjoint of the rain objective (asph_adjoint_rain).
//
// TEST INFRASTRUCTURE ONLY: the reference the GPU adjoint is checked against, beside central
// finite differences of the FP64 oracle. It is loaded by tests/ and bench.py's checks only.
//
// The forward is the FP64 instance of the oracle (asph_oracle.hpp), recorded step by step in host
// memory (no checkpoints). The reverse sweep differentiates the same stages with the same
// pointwise and pair derivative formulas as the device (asph_adjoint.h, instantiated for double)
// but with its own orchestration: plain loops over the oracle's row-major cell ranges, the
// stencil tests written out, a full tape instead of recomputation, and double everywhere. It
// pins the device's gather and bookkeeping and measures the FP32 effects; the finite
// differences pin the formulas.
#include <algorithm>
using std::max;
using std::min;

#include "asph_adjoint.h"
#include "asph_oracle.hpp"

namespace {

using oracle::Oracle;
using oracle::Params;

// The device library's derived parameters (float evaluation, as make_params).
asph::DevParams devparams(const asph_config& c) {
    asph::DevParams D;
    std::memset(&D, 0, sizeof(D));
    Params<float> P(c);
    D.nx = c.nx; D.ny = c.ny; D.nz = c.nz; D.S = P.S;
    D.N = c.nx * c.ny * c.nz; D.ncell = D.N; D.ncol = c.nx * c.ny;
    D.end_bit = oracle::sort_end_bit(c.nx, c.ny, c.nz);
    D.flags = c.flags;
    D.dx = P.dx; D.dy = P.dy; D.dz = P.dz; D.inv_dx = P.inv_dx; D.inv_dy = P.inv_dy; D.inv_dz = P.inv_dz;
    D.Lx = P.Lx; D.Ly = P.Ly; D.Lz = P.Lz; D.hLx = P.hLx; D.hLy = P.hLy;
    D.fnx = P.fnx; D.fny = P.fny; D.hnx = P.hnx; D.hny = P.hny;
    D.gmx = P.gmx; D.gmy = P.gmy; D.gmz = P.gmz;
    D.inv_h = P.inv_h; D.lim = P.lim; D.kin = P.kin; D.kout = P.kout;
    D.vcell = P.vcell; D.rho0 = P.rho0; D.pmass = P.pmass; D.mm = P.mm; D.eta2 = P.eta2; D.visc_c = P.visc_c;
    D.cs = P.cs; D.cs2 = P.cs2; D.th0 = P.th0; D.n2 = P.n2; D.exA = P.exA; D.cp_rd = P.cp_rd; D.rd_cp = P.rd_cp;
    D.sponge_sigma = P.sponge_sigma; D.sponge_frac = P.sponge_frac;
    D.sponge_zs = (1.0f - c.sponge_frac) * P.Lz; D.sponge_depth = c.sponge_frac * P.Lz;
    D.dthdt_cap = P.dthdt_cap; D.c_h = P.c_h; D.c_e = P.c_e; D.beta_moist = P.beta_moist; D.v_bg = P.v_bg;
    D.dts_skin = P.dts_skin; D.ramp = P.ramp;
    D.qc_auto = P.qc_auto; D.rate_auto = P.rate_auto; D.k_accr = P.k_accr; D.v_rain = P.v_rain;
    D.theta_cap = P.theta_cap; D.qv_cap = P.qv_cap;
    D.qv_mode = P.qv_mode; D.qv0 = P.qv0; D.qv_hs = P.qv_hs; D.qv_zt = P.qv_zt; D.qv_ht = P.qv_ht;
    D.rh_n = P.rh_n;
    for (int k = 0; k < ASPH_RH_KNOTS; ++k) { D.rh_z[k] = c.rh_z[k]; D.rh_v[k] = c.rh_v[k]; }
    return D;
}

struct Tape {
    std::vector<double> x0, v0, a0, t0, r0, x1, v1, r1, s12, t1, prs, a1, av;
    std::vector<int64_t> rain0;
    std::vector<uint32_t> keys;
    std::vector<int32_t> cell_range;
    std::vector<int32_t> cur;      // current cell (ix, iy, iz) after the drift, in the forward's arithmetic
    std::vector<int64_t> cid1;     // rain cell (cell_of) after the drift
    double fxs = 0.0;              // fixed-point scale of the forward's rain grid
    double Ts = 0.0, qs = 0.0;     // surface skin temperature and saturation of the step
    int mg[3] = {1, 1, 1};         // largest |current cell - listed cell| of the step per axis
};

struct Adj {
    std::vector<double> x, v, a, t, rho, rain;
    explicit Adj(int64_t N) : x(4 * N), v(4 * N), a(4 * N), t(4 * N), rho(N), rain(N) {}
};

template <class A, class B> void put(std::vector<A>& d, const std::vector<B>& s) { d.assign(s.begin(), s.end()); }

template <class F> void rec_pre(const Oracle<F>& o, Tape& T) {
    put(T.x0, o.pos); put(T.v0, o.vel); put(T.a0, o.acc); put(T.t0, o.thq); put(T.r0, o.rho);
    T.rain0 = o.rain;
    T.keys = o.keys_sorted; T.cell_range = o.cell_range;
    T.fxs = Oracle<F>::FXS;
}

template <class F> void rec_mid(const Oracle<F>& o, Tape& T) {
    put(T.x1, o.pos); put(T.v1, o.vel); put(T.r1, o.rho); put(T.t1, o.thq); put(T.prs, o.prs);
    const int64_t N = o.N;
    T.cur.resize(3 * (size_t)N);
    T.cid1.resize((size_t)N);
    for (int64_t i = 0; i < N; ++i) {
        o.P.cell_trunc(&o.pos[4 * i], T.cur[3 * i], T.cur[3 * i + 1], T.cur[3 * i + 2]);
        T.cid1[i] = o.P.cell_of(&o.pos[4 * i]);
    }
}

int wrapd(int d, int n) {   // periodic offset in (-n/2, n/2]
    d %= n;
    if (d < 0) d += n;
    if (d > n / 2) d -= n;
    return d;
}

bool in_stencil(const asph::DevParams& P, int ix, int iy, int iz, int cx, int cy, int cz) {
    const int S = P.S;
    const bool interior = (iz >= S && iz < P.nz - S);
    const int z0 = interior ? iz - S : std::max(iz - S, 0);
    const int z1 = interior ? iz + S : std::min(iz + S, P.nz - 1);
    if (cz < z0 || cz > z1) return false;
    return std::abs(wrapd(cx - ix, P.nx)) <= S && std::abs(wrapd(cy - iy, P.ny)) <= S;
}


// Separation a - b in cells (the forward's form: coordinates scaled to cells, then subtracted;
// minimum image by the period in cells), in T = double or, as the device, float.
template <class T> void csep(const asph::DevParams& P, const T* a, const T* b, T s[3]) {
    const T inv[3] = {(T)P.inv_dx, (T)P.inv_dy, (T)P.inv_dz};
    for (int c = 0; c < 3; ++c) s[c] = a[c] * inv[c] - b[c] * inv[c];
    if (P.flags & ASPH_F_MIN_IMAGE) {
        const T nx = (T)P.nx, ny = (T)P.ny;
        if (s[0] > T(0.5) * nx) s[0] -= nx; else if (s[0] < -T(0.5) * nx) s[0] += nx;
        if (s[1] > T(0.5) * ny) s[1] -= ny; else if (s[1] < -T(0.5) * ny) s[1] += ny;
    }
}
void sep(const asph::DevParams& P, const double* a, const double* b, double r[3]) { csep<double>(P, a, b, r); }

// Cells added to the enumeration margin of for_pairs (a test setting: the gradient must not change).
int g_extra_margin = 0;
// Reverse stages evaluated in float as the device does (a measurement setting; 0 = all double).
enum { FS_MOM_PAIRS = 1, FS_DENS_PAIRS = 2, FS_EOS = 4, FS_STATE = 8, FS_MOM_EPI = 16, FS_DENS_RATIO = 32 };
int g_float_stages = 0;

// Largest offset per axis between a particle's current cell (after the drift) and its listed cell.
void drift_margin(const asph::DevParams& P, Tape& T) {
    T.mg[0] = T.mg[1] = T.mg[2] = 0;
    for (int64_t i = 0; i < P.N; ++i) {
        const uint32_t key = T.keys[(size_t)i];
        const int* c = &T.cur[3 * (size_t)i];
        const int l[3] = {(int)asph::compact1by2(key), (int)asph::compact1by2(key >> 1),
                          (int)asph::compact1by2(key >> 2)};
        T.mg[0] = std::max(T.mg[0], std::abs(wrapd(c[0] - l[0], P.nx)));
        T.mg[1] = std::max(T.mg[1], std::abs(wrapd(c[1] - l[1], P.ny)));
        T.mg[2] = std::max(T.mg[2], std::abs(c[2] - l[2]));
    }
}

// Calls fn(j, own, rev) for the particles j for which (i, j) is a pair of i's sum (own: j listed in
// the stencil of i's current cell) or (j, i) one of j's sum (rev: i listed in the stencil of j's
// current cell). Both kinds are listed within S + mg cells of i's listed cell per axis (mg: the
// step's largest current-to-listed offset); every cell in that range is visited once.
template <class Fn>
void for_pairs(const asph::DevParams& P, const Tape& T, int64_t i, Fn fn) {
    const uint32_t key = T.keys[i];
    const int lx = (int)asph::compact1by2(key), ly = (int)asph::compact1by2(key >> 1),
              lz = (int)asph::compact1by2(key >> 2);
    const int ix = T.cur[3 * i], iy = T.cur[3 * i + 1], iz = T.cur[3 * i + 2];
    const int Mx = P.S + T.mg[0] + g_extra_margin, My = P.S + T.mg[1] + g_extra_margin,
              Mz = P.S + T.mg[2] + g_extra_margin;
    const int nxs = std::min(2 * Mx + 1, P.nx), nys = std::min(2 * My + 1, P.ny);
    for (int cz = std::max(lz - Mz, 0); cz <= std::min(lz + Mz, P.nz - 1); ++cz)
        for (int oy = 0; oy < nys; ++oy)
            for (int ox = 0; ox < nxs; ++ox) {
                const int cx = Params<double>::wrap(lx - Mx + ox, P.nx), cy = Params<double>::wrap(ly - My + oy, P.ny);
                const int64_t c = ((int64_t)cz * P.ny + cy) * P.nx + cx;
                const bool own = in_stencil(P, ix, iy, iz, cx, cy, cz);
                for (int32_t j = T.cell_range[2 * c]; j < T.cell_range[2 * c + 1]; ++j) {
                    const int* cj = &T.cur[3 * (int64_t)j];
                    const int jx = cj[0], jy = cj[1], jz = cj[2];
                    const bool rev = in_stencil(P, jx, jy, jz, lx, ly, lz);
                    if (own || rev) fn((int64_t)j, own, rev);
                }
            }
}

// Reverse of one step (mode 1) or of the force evaluation (mode 0): E (adjoints after) -> S.
// Evaluates a function of float arguments at double values rounded to float.
template <class R, int K> struct Cast {
    R v[K];
    template <class Src> explicit Cast(const Src* s) { for (int k = 0; k < K; ++k) v[k] = (R)s[k]; }
};

void reverse(const asph::DevParams& P, const Params<double>& Q, const Tape& T, const Adj& E, Adj& S, double dt,
             int mode, double auto_f, double& qc0b) {
    const bool f_mom = g_float_stages & FS_MOM_PAIRS, f_dens = g_float_stages & FS_DENS_PAIRS,
               f_eos = g_float_stages & FS_EOS, f_state = g_float_stages & FS_STATE,
               f_epi = g_float_stages & FS_MOM_EPI, f_ratio = g_float_stages & FS_DENS_RATIO;
    const int64_t N = P.N, plane = (int64_t)P.nx * P.ny;
    const bool sed = mode == 1 && (P.flags & ASPH_F_SEDIMENT);
    const double frac = sed ? (double)asph::rain_frac(P, (float)dt) : 0.0;
    std::vector<double> rb_cur(N);
    for (int64_t c = 0; c < N; ++c)
        rb_cur[c] = sed ? (1.0 - frac) * E.rain[c] + (c >= plane ? frac * E.rain[c - plane] : 0.0) : E.rain[c];
    std::vector<double> Ab(4 * N), AVb(4 * N), x1b(4 * N), v1b(4 * N), t2b(4 * N), pb(N), r1b(N), sb(2 * N);
    const double mm = P.mm;
    // momentum epilogue
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        const double* x1 = &T.x1[4 * i];
        const double* t1 = &T.t1[4 * i];
        double in[asph::adj::MOM_IN], ob[asph::adj::MOM_OUT], ib[asph::adj::MOM_IN];
        const double thv = t1[0] * (1.0 + 0.608 * t1[1]);
        const double thvr = Q.theta_ref(x1[2]) * (1.0 + 0.608 * Q.qv_ref(x1[2]));
        const double buoy = Q.GRAV * ((thv - thvr) / thvr - t1[2]);
        in[0] = T.a1[4 * i]; in[1] = T.a1[4 * i + 1]; in[2] = T.a1[4 * i + 2] - buoy;
        for (int a = 0; a < 3; ++a) { in[3 + a] = T.av[4 * i + a]; in[6 + a] = T.v1[4 * i + a]; }
        in[9] = x1[2]; in[10] = t1[0]; in[11] = t1[1]; in[12] = t1[2]; in[13] = T.prs[i];
        ob[0] = E.a[4 * i]; ob[1] = E.a[4 * i + 1]; ob[2] = E.a[4 * i + 2]; ob[3] = E.t[4 * i];
        ob[4] = E.v[4 * i]; ob[5] = E.v[4 * i + 1]; ob[6] = E.v[4 * i + 2];
        if (f_epi) {
            Cast<float, asph::adj::MOM_IN> inf(in);
            Cast<float, asph::adj::MOM_OUT> obf(ob);
            float ibf[asph::adj::MOM_IN];
            asph::adj::mom_vjp_rev<float>(P, (float)dt, mode, inf.v, obf.v, ibf);
            for (int k = 0; k < asph::adj::MOM_IN; ++k) ib[k] = ibf[k];
        } else {
            asph::adj::mom_vjp<double>(P, dt, mode, in, ob, ib);
        }
        for (int a = 0; a < 3; ++a) {
            Ab[4 * i + a] = ib[a] * mm;
            AVb[4 * i + a] = ib[3 + a] * mm;
            x1b[4 * i + a] = E.x[4 * i + a];
            v1b[4 * i + a] = ib[6 + a];
        }
        x1b[4 * i + 2] += ib[9];
        t2b[4 * i] = ib[10];
        t2b[4 * i + 1] = E.t[4 * i + 1] + ib[11];
        t2b[4 * i + 2] = E.t[4 * i + 2] + ib[12];
        pb[i] = ib[13];
        r1b[i] = E.rho[i];
    }
    // momentum pairs
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < N; ++i) {
        const double* xi = &T.x1[4 * i];
        const double vi[3] = {T.v1[4 * i], T.v1[4 * i + 1], T.v1[4 * i + 2]};
        const double rho_i = std::max(T.r1[i], 1e-3), p_i = T.prs[i];
        double xb[3] = {0, 0, 0}, vb[3] = {0, 0, 0}, rb = 0, pbi = 0;
        if (f_mom) {   // the device's arithmetic: float throughout, float accumulators
            float xf[3] = {0, 0, 0}, vf[3] = {0, 0, 0}, rf = 0, pf = 0;
            const Cast<float, 3> vif(vi), Aif(&Ab[4 * i]), AVif(&AVb[4 * i]);
            const float rhoif = std::max((float)T.r1[i], 1e-3f), pif = (float)T.prs[i];
            for_pairs(P, T, i, [&](int64_t j, bool own, bool rev) {
                const Cast<float, 4> xif(xi), xjf(&T.x1[4 * j]);
                float r[3];
                csep<float>(P, xif.v, xjf.v, r);
                const Cast<float, 3> vjf(&T.v1[4 * j]);
                const float rhojf = std::max((float)T.r1[j], 1e-3f), pjf = (float)T.prs[j];
                const float z3[3] = {0, 0, 0};
                const Cast<float, 3> Ajf(&Ab[4 * j]), AVjf(&AVb[4 * j]);
                asph::adj::mom_pair_vjp2<float>(P, r[0], r[1], r[2], vif.v, vjf.v, rhoif, rhojf, pif, pjf,
                                                own ? Aif.v : z3, own ? AVif.v : z3, rev ? Ajf.v : z3,
                                                rev ? AVjf.v : z3, xf, vf, rf, pf);
            });
            for (int a = 0; a < 3; ++a) { xb[a] = xf[a]; vb[a] = vf[a]; }
            rb = rf;
            pbi = pf;
        } else
        for_pairs(P, T, i, [&](int64_t j, bool own, bool rev) {
            const double* xj = &T.x1[4 * j];
            const double vj[3] = {T.v1[4 * j], T.v1[4 * j + 1], T.v1[4 * j + 2]};
            const double rho_j = std::max(T.r1[j], 1e-3), p_j = T.prs[j];
            asph::adj::PairBar<double> b;
            double r[3];
            if (own) {
                sep(P, xi, xj, r);
                if (asph::adj::mom_pair_vjp<double>(P, r[0], r[1], r[2], vi, vj, rho_i, rho_j, p_i, p_j,
                                                    &Ab[4 * i], &AVb[4 * i], b)) {
                    for (int a = 0; a < 3; ++a) { xb[a] += b.r[a]; vb[a] += b.vd[a]; }
                    rb += b.rho_o;
                    pbi += b.p_o;
                }
            }
            if (rev) {
                sep(P, xj, xi, r);
                if (asph::adj::mom_pair_vjp<double>(P, r[0], r[1], r[2], vj, vi, rho_j, rho_i, p_j, p_i,
                                                    &Ab[4 * j], &AVb[4 * j], b)) {
                    for (int a = 0; a < 3; ++a) { xb[a] -= b.r[a]; vb[a] -= b.vd[a]; }
                    rb += b.rho_n;
                    pbi += b.p_n;
                }
            }
        });
        for (int a = 0; a < 3; ++a) { x1b[4 * i + a] += xb[a]; v1b[4 * i + a] += vb[a]; }
        if (T.r1[i] > 1e-3) r1b[i] += rb;
        pb[i] += pbi;
    }
    // EOS + microphysics
    S.rain = rb_cur;
    const double Ts = T.Ts, qs = T.qs;
    for (int64_t i = 0; i < N; ++i) {
        const double* x1 = &T.x1[4 * i];
        const int64_t cid = T.cid1[i];
        asph::adj::EosConst<double> K;
        K.dt = dt;
        K.auto_f = auto_f;   // the forward's own constant
        K.Ts = Ts;
        K.qs_sfc = qs;
        K.flux = mode == 1 && (P.flags & ASPH_F_SURFACE_FLUX) && (T.keys[i] & 0x24924924u) == 0u && x1[2] < (double)P.dz;
        const double in[asph::adj::EOS_IN] = {T.t0[4 * i], T.t0[4 * i + 1], T.t0[4 * i + 2], T.r1[i], x1[2],
                                              (double)T.rain0[cid] / T.fxs, (double)P.qc_auto};
        const double ob[asph::adj::EOS_OUT] = {t2b[4 * i], t2b[4 * i + 1], t2b[4 * i + 2], pb[i], rb_cur[cid]};
        double ib[asph::adj::EOS_IN];
        if (f_eos) {
            asph::adj::EosConst<float> Kf;
            Kf.dt = (float)dt; Kf.auto_f = (float)auto_f; Kf.Ts = (float)Ts; Kf.qs_sfc = (float)qs; Kf.flux = K.flux;
            Cast<float, asph::adj::EOS_IN> inf(in);
            Cast<float, asph::adj::EOS_OUT> obf(ob);
            float ibf[asph::adj::EOS_IN];
            asph::adj::eos_vjp<float>(P, Kf, inf.v, obf.v, ibf);
            for (int k = 0; k < asph::adj::EOS_IN; ++k) ib[k] = ibf[k];
        } else {
            asph::adj::eos_vjp<double>(P, K, in, ob, ib);
        }
        S.t[4 * i] = ib[0]; S.t[4 * i + 1] = ib[1]; S.t[4 * i + 2] = ib[2]; S.t[4 * i + 3] = 0;
        r1b[i] += ib[3];
        x1b[4 * i + 2] += ib[4];
        S.rain[cid] += ib[5];
        qc0b += ib[6];
    }
    // density ratio
    for (int64_t i = 0; i < N; ++i) {
        double zb;
        if (f_ratio) {
            float s1b, s2b, zbf;
            asph::adj::dens_vjp<float>(P, (float)T.s12[2 * i], (float)T.s12[2 * i + 1], (float)T.x1[4 * i + 2],
                                       (float)r1b[i], s1b, s2b, zbf);
            sb[2 * i] = s1b; sb[2 * i + 1] = s2b; zb = zbf;
        } else {
            asph::adj::dens_vjp<double>(P, T.s12[2 * i], T.s12[2 * i + 1], T.x1[4 * i + 2], r1b[i], sb[2 * i],
                                        sb[2 * i + 1], zb);
        }
        x1b[4 * i + 2] += zb;
    }
    // density pairs
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < N; ++i) {
        const double* xi = &T.x1[4 * i];
        const double inv_r0i = 1.0 / std::max(T.r0[i], 1e-3);
        double xb[3] = {0, 0, 0}, rb = 0;
        if (f_dens) {
            float xf[3] = {0, 0, 0}, rf = 0;
            const float ir0f = 1.0f / std::max((float)T.r0[i], 1e-3f);
            for_pairs(P, T, i, [&](int64_t j, bool own, bool rev) {
                const Cast<float, 4> xif(xi), xjf(&T.x1[4 * j]);
                float r[3];
                csep<float>(P, xif.v, xjf.v, r);
                float W, gw[3];
                if (!asph::adj::dens_pair<float>(P, r[0], r[1], r[2], W, gw)) return;
                float wb = 0.f;
                if (own) wb = (float)sb[2 * i] + (float)sb[2 * i + 1] * (1.0f / std::max((float)T.r0[j], 1e-3f));
                if (rev) {
                    wb += (float)sb[2 * j] + (float)sb[2 * j + 1] * ir0f;
                    rf += (float)sb[2 * j + 1] * W * (-ir0f * ir0f);
                }
                for (int a = 0; a < 3; ++a) xf[a] += wb * gw[a];
            });
            for (int a = 0; a < 3; ++a) xb[a] = xf[a];
            rb = rf;
        } else
        for_pairs(P, T, i, [&](int64_t j, bool own, bool rev) {
            const double* xj = &T.x1[4 * j];
            double r[3], W, gw[3];
            if (own) {
                sep(P, xi, xj, r);
                if (asph::adj::dens_pair<double>(P, r[0], r[1], r[2], W, gw)) {
                    const double wb = sb[2 * i] + sb[2 * i + 1] * (1.0 / std::max(T.r0[j], 1e-3));
                    for (int a = 0; a < 3; ++a) xb[a] += wb * gw[a];
                }
            }
            if (rev) {
                sep(P, xj, xi, r);
                if (asph::adj::dens_pair<double>(P, r[0], r[1], r[2], W, gw)) {
                    const double wb = sb[2 * j] + sb[2 * j + 1] * inv_r0i;
                    for (int a = 0; a < 3; ++a) xb[a] -= wb * gw[a];
                    rb += sb[2 * j + 1] * W * (-inv_r0i * inv_r0i);
                }
            }
        });
        for (int a = 0; a < 3; ++a) x1b[4 * i + a] += xb[a];
        S.rho[i] = T.r0[i] > 1e-3 ? rb : 0.0;
    }
    // kick + drift
    for (int64_t i = 0; i < N; ++i) {
        double* sx = &S.x[4 * i];
        double* sv = &S.v[4 * i];
        double* sa = &S.a[4 * i];
        const double* xb = &x1b[4 * i];
        const double* vb = &v1b[4 * i];
        if (mode == 0) {
            for (int a = 0; a < 3; ++a) { sx[a] = xb[a]; sv[a] = vb[a]; sa[a] = 0; }
            sx[3] = sv[3] = sa[3] = 0;
            continue;
        }
        const double hdt = 0.5 * dt;
        const double vz = std::fma(T.a0[4 * i + 2], hdt, T.v0[4 * i + 2]);
        const double z1 = std::fma(vz, dt, T.x0[4 * i + 2]);
        const double sg = (z1 < 0.0 || z1 > (double)Q.Lz) ? -1.0 : 1.0;
        const double zb = sg * xb[2];
        const double vp[3] = {vb[0] + dt * xb[0], vb[1] + dt * xb[1], sg * vb[2] + dt * zb};
        sx[0] = xb[0]; sx[1] = xb[1]; sx[2] = zb; sx[3] = 0;
        for (int a = 0; a < 3; ++a) { sv[a] = vp[a]; sa[a] = hdt * vp[a]; }
        sv[3] = sa[3] = 0;
    }
    if (f_state) {   // the device keeps the state adjoints in float
        for (auto* v : {&S.x, &S.v, &S.a, &S.t, &S.rho})
            for (double& x : *v) x = (double)(float)x;
    }
}

void unperm(const std::vector<uint32_t>& perm, Adj& A) {
    const int64_t N = (int64_t)perm.size();
    Adj B(N);
    for (int64_t i = 0; i < N; ++i) {
        const int64_t k = perm[i];
        for (int a = 0; a < 4; ++a) {
            B.x[4 * k + a] = A.x[4 * i + a];
            B.v[4 * k + a] = A.v[4 * i + a];
            B.a[4 * k + a] = A.a[4 * i + a];
            B.t[4 * k + a] = A.t[4 * i + a];
        }
        B.rho[k] = A.rho[i];
    }
    B.rain = A.rain;
    A = std::move(B);
}

// The forward in F (the FP64 or the FP32 oracle), recorded into a double tape; the reverse in double.
template <class F>
int adjoint_impl(const asph_config* cfg, const double* pos4, const double* vel4, const double* thq4,
                 const double* rho, int64_t nsteps, float dt_f, int ck, double* J_out, double* g_thq4,
                 double* g_qc0) {
    if (!cfg || !pos4 || !vel4 || !thq4 || !rho || nsteps < 1 || ck < 1) return 1;
    const asph::DevParams P = devparams(*cfg);
    Oracle<F> o(*cfg);
    const int64_t N = o.N;
    for (int64_t k = 0; k < 4 * N; ++k) {
        o.pos[k] = (F)pos4[k];
        o.vel[k] = (F)vel4[k];
        o.thq[k] = (F)thq4[k];
    }
    for (int64_t k = 0; k < N; ++k) o.rho[k] = (F)rho[k];
    const double dt = (double)dt_f;
    const Params<double> Q(*cfg);   // the reverse's constants, in double
    // the forward's autoconversion factor (1 - exp(-k dt)): in float it carries the cancellation
    const double auto_f = (double)(F(1) - oracle::mexp<F>(-o.P.rate_auto * (F)dt_f));
    std::vector<std::vector<uint32_t>> perms;
    std::vector<Tape> tape((size_t)nsteps + 1);   // slot 0: the force evaluation
    std::vector<F> s12(2 * N), av(4 * N);
    // forward
    o.sort();
    perms.push_back(o.perm);
    {
        Tape& T = tape[0];
        rec_pre(o, T);
        std::fill(o.dep.begin(), o.dep.end(), 0);
        o.rec_s12 = &s12;
        o.density();
        o.rec_s12 = nullptr;
        o.eos_micro(dt);
        rec_mid(o, T);
        drift_margin(P, T);
        put(T.s12, s12);
        o.rain_merge();
        o.rec_av = &av;
        o.momentum();
        o.rec_av = nullptr;
        put(T.a1, o.acc);
        put(T.av, av);
        o.initialized = true;
    }
    for (int64_t k = 0; k < nsteps; ++k) {
        const double t = (double)k * dt;
        if (k > 0 && k % ck == 0) {
            o.sort();
            perms.push_back(o.perm);
        }
        Tape& T = tape[(size_t)k + 1];
        {
            F ts, qs;
            o.surface_flux_params(t, ts, qs);
            T.Ts = (double)ts;
            T.qs = (double)qs;
        }
        rec_pre(o, T);
        std::fill(o.dep.begin(), o.dep.end(), 0);
        o.kick(0.5 * dt);
        o.drift(dt);
        if (o.flag(ASPH_F_SURFACE_FLUX)) o.surface_flux(dt, t);
        o.rec_s12 = &s12;
        o.density();
        o.rec_s12 = nullptr;
        o.eos_micro(dt);
        // x1 and v1 are the positions and velocities after the drift; the flux changed thq only
        rec_mid(o, T);
        drift_margin(P, T);
        put(T.s12, s12);
        o.rec_av = &av;
        o.momentum();
        o.rec_av = nullptr;
        put(T.a1, o.acc);
        put(T.av, av);
        o.theta_update(dt);
        o.kick(0.5 * dt);
        if (o.flag(ASPH_F_SPONGE)) o.sponge(dt);
        o.rain_adv(dt);
    }
    long double sum = 0;
    for (int64_t c = 0; c < N; ++c) sum += (long double)o.rain[c];
    *J_out = (double)(sum / (long double)Oracle<F>::FXS) * (double)P.pmass;
    // reverse
    Adj E(N), S(N);
    std::fill(E.rain.begin(), E.rain.end(), (double)P.pmass);
    double qc0b = 0.0;
    for (int64_t k = nsteps - 1; k >= 0; --k) {
        reverse(P, Q, tape[(size_t)k + 1], E, S, dt, 1, auto_f, qc0b);
        std::swap(E, S);
        if (k > 0 && k % ck == 0) unperm(perms[(size_t)(k / ck)], E);
    }
    reverse(P, Q, tape[0], E, S, dt, 0, auto_f, qc0b);
    std::swap(E, S);
    unperm(perms[0], E);
    std::memcpy(g_thq4, E.t.data(), 4 * N * sizeof(double));
    *g_qc0 = qc0b;
    return 0;
}


}  // namespace

extern "C" {

void oracle_adjoint_set_extra_margin(int cells) { g_extra_margin = cells < 0 ? 0 : cells; }
void oracle_adjoint_set_float_stages(int mask) { g_float_stages = mask; }

// The fused two-orientation momentum pair (mom_pair_vjp2) against two calls of mom_pair_vjp on
// random pairs of the configuration's support: returns the largest difference relative to the
// largest term, in double (mode 0) or float (mode 1).
// mom_vjp_rev against the forward-mode mom_vjp on random inputs around the configuration's state
// (including active clamps and the sponge layer): the largest difference relative to the largest
// adjoint, in double (prec 0) or float (prec 1), for the step (mode 1) or the force evaluation (0).
double oracle_check_mom_vjp_rev(const asph_config* cfg, int n, uint64_t seed, int prec, int mode) {
    const asph::DevParams P = devparams(*cfg);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 3;
    auto rnd = [&]() {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);   // [0, 1)
    };
    double worst = 0.0;
    auto run = [&](auto zero) {
        using R = decltype(zero);
        for (int k = 0; k < n; ++k) {
            R in[asph::adj::MOM_IN], ob[asph::adj::MOM_OUT], g1[asph::adj::MOM_IN], g2[asph::adj::MOM_IN];
            for (int a = 0; a < 9; ++a) in[a] = (R)(20 * (rnd() - 0.5));
            if (k % 5 == 0) for (int a = 3; a < 6; ++a) in[a] = (R)(2e4 * (rnd() - 0.5));   // theta clamps
            in[9] = (R)(rnd() * P.Lz);
            in[10] = (R)(280 + 40 * rnd());
            in[11] = (R)(0.015 * rnd());
            in[12] = (R)(0.002 * rnd());
            in[13] = (R)(200 * (rnd() - 0.5));
            if (k % 7 == 0) in[13] = (R)(-2e6);   // pressure floor
            if (k % 11 == 0) in[10] = (R)(P.theta_cap + 0.5);   // theta clamps of the update
            if (k % 13 == 0) in[10] = (R)99.5;
            for (int a = 0; a < asph::adj::MOM_OUT; ++a) ob[a] = (R)(2 * rnd() - 1);
            const R dtr = (R)(0.25 + rnd());
            asph::adj::mom_vjp<R>(P, dtr, mode, in, ob, g1);
            asph::adj::mom_vjp_rev<R>(P, dtr, mode, in, ob, g2);
            double big = 0, diff = 0;
            for (int a = 0; a < asph::adj::MOM_IN; ++a) {
                big = std::max(big, std::fabs((double)g1[a]));
                diff = std::max(diff, std::fabs((double)g1[a] - (double)g2[a]));
            }
            if (big > 0) worst = std::max(worst, diff / big);
        }
    };
    if (prec == 0) run(0.0); else run(0.0f);
    return worst;
}

double oracle_check_mom_pair_fused(const asph_config* cfg, int n, uint64_t seed, int mode) {
    const asph::DevParams P = devparams(*cfg);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto rnd = [&]() {   // uniform in [-1, 1)
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        return (double)(st >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    };
    double worst = 0.0;
    bool mismatch = false;
    auto run = [&](auto zero) {
        using R = decltype(zero);
        for (int k = 0; k < n; ++k) {
            const R R0 = (R)std::sqrt((double)P.lim) * (R)0.95;
            R r[3];
            r[0] = (R)(rnd() * R0 * 0.577);   // separations in cells
            r[1] = (R)(rnd() * R0 * 0.577);
            r[2] = (R)(rnd() * R0 * 0.577);
            R vi[3], vj[3], Ai[3], AVi[3], Aj[3], AVj[3];
            for (int a = 0; a < 3; ++a) {
                vi[a] = (R)(5 * rnd()); vj[a] = (R)(5 * rnd());
                Ai[a] = (R)rnd(); AVi[a] = (R)rnd(); Aj[a] = (R)rnd(); AVj[a] = (R)rnd();
            }
            if (k % 4 == 1) for (int a = 0; a < 3; ++a) Ai[a] = AVi[a] = 0;   // only (j, i)
            if (k % 4 == 2) for (int a = 0; a < 3; ++a) Aj[a] = AVj[a] = 0;   // only (i, j)
            const R rho_i = (R)(1 + 0.2 * rnd()), rho_j = (R)(1 + 0.2 * rnd());
            const R p_i = (R)(50 * rnd()), p_j = (R)(50 * rnd());
            R xa[3] = {0, 0, 0}, va[3] = {0, 0, 0}, ra = 0, pa = 0;
            asph::adj::PairBar<R> b;
            bool in1 = asph::adj::mom_pair_vjp<R>(P, r[0], r[1], r[2], vi, vj, rho_i, rho_j, p_i, p_j, Ai, AVi, b);
            if (in1) {
                for (int a = 0; a < 3; ++a) { xa[a] += b.r[a]; va[a] += b.vd[a]; }
                ra += b.rho_o; pa += b.p_o;
            }
            bool in2 = asph::adj::mom_pair_vjp<R>(P, -r[0], -r[1], -r[2], vj, vi, rho_j, rho_i, p_j, p_i, Aj, AVj, b);
            if (in2) {
                for (int a = 0; a < 3; ++a) { xa[a] -= b.r[a]; va[a] -= b.vd[a]; }
                ra += b.rho_n; pa += b.p_n;
            }
            R xf[3] = {0, 0, 0}, vf[3] = {0, 0, 0}, rf = 0, pf = 0;
            bool in3 = asph::adj::mom_pair_vjp2<R>(P, r[0], r[1], r[2], vi, vj, rho_i, rho_j, p_i, p_j, Ai, AVi, Aj,
                                                   AVj, xf, vf, rf, pf);
            if (in1 != in3 || in2 != in3) { mismatch = true; return; }
            const double ref[8] = {xa[0], xa[1], xa[2], va[0], va[1], va[2], ra, pa};
            const double got[8] = {xf[0], xf[1], xf[2], vf[0], vf[1], vf[2], rf, pf};
            double big = 0, diff = 0;
            for (int m = 0; m < 8; ++m) {
                big = std::max(big, std::fabs(ref[m]));
                diff = std::max(diff, std::fabs(ref[m] - got[m]));
            }
            if (big > 0) worst = std::max(worst, diff / big);
        }
    };
    if (mode == 0) run(0.0); else run(0.0f);
    return mismatch ? 1e30 : worst;   // 1e30: a support decision differed
}


// FP64 host adjoint of asph_adjoint_rain (same schedule and objective). Inputs are the initial
// state in the layout of asph_set_state (float64); outputs J, dJ/d(theta, qv, qc, -) per particle
// in that order, and dJ/d(qc_auto).
int oracle_adjoint_rain(const asph_config* cfg, const double* pos4, const double* vel4, const double* thq4,
                        const double* rho, int64_t nsteps, float dt_f, int ck, double* J_out, double* g_thq4,
                        double* g_qc0) {
    return adjoint_impl<double>(cfg, pos4, vel4, thq4, rho, nsteps, dt_f, ck, J_out, g_thq4, g_qc0);
}

// The same reverse sweep (double) along the FP32 oracle's trajectory, which is the device's bit
// for bit: separates the FP32 reverse arithmetic of the device from the FP32 trajectory.
int oracle_adjoint_rain_f32_trajectory(const asph_config* cfg, const double* pos4, const double* vel4,
                                       const double* thq4, const double* rho, int64_t nsteps, float dt_f, int ck,
                                       double* J_out, double* g_thq4, double* g_qc0) {
    return adjoint_impl<float>(cfg, pos4, vel4, thq4, rho, nsteps, dt_f, ck, J_out, g_thq4, g_qc0);
}

}  // extern "C"
