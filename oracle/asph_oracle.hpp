// asph_oracle.hpp - CPU restatement of the reference's anisotropic-SPH step (asph_h100.cu).
//
// TEST INFRASTRUCTURE ONLY. This is the parity oracle for the HIP implementation in
// sph_cuda_amd/. It is imported by tests/, by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg as the checker or as the CPU baseline, and never by the product path.
//
// The oracle follows the reference kernels one by one, in the reference's loop order
// (asph_h100.cu:958-1078), templated on the real type R:
//   R = float : the FP32 oracle. Its arithmetic is the project's canonical form of the
//               reference formulas (explicit fma where the device kernels use fma, exp/log/pow
//               from asph_fmath.h), so that device results can be compared bit for bit;
//   R = double: the FP64 instance, used as the finite-difference reference for the adjoint.
// Reformulations that change only the rounding of the reference's expressions (not the
// mathematics) are marked "canonical:" next to the reference line they restate.
//
// Parity pins (tests/test_oracle_*.py): Morton known answers, stable-argsort closed form,
// brute-force cell ranges and neighbour sums, spline normalisation/continuity/derivative,
// hydrostatic sounding identities, Magnus values, saturation-adjustment invariants, rain mass
// balance, KDK free-particle and reflection cases, and a straightforward float64 Python
// restatement on small lattices.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <type_traits>
#include <vector>

#include "asph.h"
#include "asph_fmath.h"

namespace oracle {

// Physical constants, asph_h100.cu:56, 66-70.
constexpr float GRAV_F = 9.80665f;
constexpr float R_D_F = 287.05f;
constexpr float CP_F = 1004.0f;
constexpr float P0_F = 100000.0f;
constexpr float L_V_F = 2.501e6f;
constexpr float SIGMA_F = 0.31830988618f;
constexpr double FX_SCALE = 1099511627776.0;  // 2^40: fixed-point rain units per particle mass

template <class R> inline R mexp(R x);
template <> inline float mexp<float>(float x) { return asphm::fexp(x); }
template <> inline double mexp<double>(double x) { return std::exp(x); }
template <class R> inline R mpow(R x, R y);
template <> inline float mpow<float>(float x, float y) { return asphm::fpow(x, y); }
template <> inline double mpow<double>(double x, double y) { return std::pow(x, y); }
template <class R> inline R mfma(R a, R b, R c);
template <> inline float mfma<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> inline double mfma<double>(double a, double b, double c) { return std::fma(a, b, c); }
template <class R> inline R msqrt(R x) { return std::sqrt(x); }
template <class R> inline R mmin(R a, R b) { return std::fmin(a, b); }
template <class R> inline R mmax(R a, R b) { return std::fmax(a, b); }

inline int64_t to_fx(double v) { return (int64_t)std::llrint(v * FX_SCALE); }

// 10-bit spread, asph_h100.cu:176-183 (bit i of n moves to bit 3i).
inline uint32_t part1by2(uint32_t n) {
    uint32_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((n >> b) & 1u) << (3 * b);
    return r;
}
// asph_h100.cu:184-186.
inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2);
}
// asph_h100.cu:866-868 (ceil_log2 starts at 1).
inline int sort_end_bit(int nx, int ny, int nz) {
    auto cl2 = [](int n) { int b = 1; while ((1 << b) < n) ++b; return b; };
    return 3 * std::max(cl2(nx), std::max(cl2(ny), cl2(nz)));
}

// Derived constants of a configuration, evaluated in R exactly as the device library does.
template <class R> struct Params {
    int nx, ny, nz, S;
    int64_t N;
    uint32_t flags;
    R dx, dy, dz, inv_dx, inv_dy, inv_dz, Lx, Ly, Lz, hLx, hLy;
    R fnx, fny, hnx, hny;   // periods in x and y [cells] and their halves
    R h, inv_h, lim, kin, kout;
    R gmx, gmy, gmz;        // mm / dx, mm / dy, mm / dz: the momentum sums' factors
    R vcell, rho0, pmass, mm, eta2, visc_c, cs, cs2;
    R th0, n2, exA, cp_rd, rd_cp;
    R sponge_sigma, sponge_frac, dthdt_cap, c_h, c_e, beta_moist, v_bg, dts_skin, ramp;
    R qc_auto, rate_auto, k_accr, v_rain, theta_cap, qv_cap;
    int qv_mode;
    R qv0, qv_hs, qv_zt, qv_ht;
    int rh_n;
    R rh_z[ASPH_RH_KNOTS], rh_v[ASPH_RH_KNOTS];
    R GRAV, R_D, CP, P0, L_V, SIGMA;

    explicit Params(const asph_config& c) {
        GRAV = R(GRAV_F); R_D = R(R_D_F); CP = R(CP_F); P0 = R(P0_F); L_V = R(L_V_F);
        SIGMA = R(SIGMA_F);
        nx = c.nx; ny = c.ny; nz = c.nz;
        N = (int64_t)nx * ny * nz;
        flags = c.flags;
        h = R(c.h);
        S = c.stencil > 0 ? c.stencil : (int)std::ceil((double)c.h * 2.0 - 1e-6);
        dx = R(c.dx); dy = R(c.dy); dz = R(c.dz);
        inv_dx = R(1) / dx; inv_dy = R(1) / dy; inv_dz = R(1) / dz;   // asph_h100.cu:49
        Lx = R(nx) * dx; Ly = R(ny) * dy; Lz = R(nz) * dz;            // :792
        hLx = R(0.5) * Lx; hLy = R(0.5) * Ly;
        fnx = R(nx); fny = R(ny);
        hnx = R(0.5) * fnx; hny = R(0.5) * fny;
        inv_h = R(1) / h;
        lim = (R(4) * h) * h;                                          // :319, :457
        R ih5 = (((inv_h * inv_h) * inv_h) * inv_h) * inv_h;
        kin = SIGMA * ih5;
        kout = (R(-0.75) * SIGMA) * ih5;
        vcell = (dx * dy) * dz;                                        // :50
        th0 = R(c.th0); n2 = R(c.n2); cs = R(c.cs);
        cs2 = cs * cs;
        rho0 = P0 / (R_D * th0);                                       // :75
        pmass = rho0 * vcell;                                          // :76
        mm = pmass / vcell;                                            // :497
        gmx = mm * inv_dx; gmy = mm * inv_dy; gmz = mm * inv_dz;       // canonical :463-465, :497
        R hav = R(c.h_av);
        eta2 = (R(0.01) * hav) * (R(0.01) * hav);                      // :482
        visc_c = (-R(c.alpha_av) * cs) * hav;                          // :486
        exA = n2 > R(0) ? (GRAV * GRAV) / ((CP * th0) * n2) : R(0);    // :140
        cp_rd = CP / R_D; rd_cp = R_D / CP;
        sponge_sigma = R(c.sponge_sigma); sponge_frac = R(c.sponge_frac);
        dthdt_cap = R(c.dthdt_cap);
        c_h = R(c.c_h); c_e = R(c.c_e); beta_moist = R(c.beta_moist); v_bg = R(c.v_bg);
        dts_skin = R(c.dts_skin); ramp = R(c.ramp);
        qc_auto = R(c.qc_auto); rate_auto = R(c.rate_auto); k_accr = R(c.k_accr);
        v_rain = R(c.v_rain); theta_cap = R(c.theta_cap); qv_cap = R(c.qv_cap);
        qv_mode = c.qv_mode;
        qv0 = R(c.qv0); qv_hs = R(c.qv_hscale); qv_zt = R(c.qv_ztop); qv_ht = R(c.qv_htop);
        rh_n = c.rh_n;
        for (int k = 0; k < ASPH_RH_KNOTS; ++k) { rh_z[k] = R(c.rh_z[k]); rh_v[k] = R(c.rh_v[k]); }
    }

    // --- reference state, asph_h100.cu:136-153 ---------------------------------------------
    R theta_ref(R z) const {
        if (n2 > R(0)) return th0 * mexp<R>((n2 * z) / GRAV);
        return th0;
    }
    R exner_ref(R z) const {
        if (n2 > R(0)) return R(1) - exA * (R(1) - mexp<R>((-n2 * z) / GRAV));
        return R(1) - (GRAV * z) / (CP * th0);  // N^2 -> 0 limit of :141
    }
    R pres_ref(R z) const { return P0 * mpow<R>(exner_ref(z), cp_rd); }
    R rho_ref(R z) const { return pres_ref(z) / ((R_D * theta_ref(z)) * exner_ref(z)); }
    // --- thermodynamics, asph_h100.cu:155-171 ------------------------------------------------
    R esat(R T) const {
        R Tc = T - R(273.15f);
        return R(611.2f) * mexp<R>((R(17.67f) * Tc) / (Tc + R(243.5f)));
    }
    R qsat(R T, R p) const {
        R es = mmin<R>(esat(T), R(0.5f) * p);
        R den = p - R(0.378f) * es;
        return den > R(1) ? (R(0.622f) * es) / den : R(0);
    }
    R dqsdT(R T, R es, R p) const {
        R Tc = T - R(273.15f);
        R d = Tc + R(243.5f);
        R desdT = (es * R(17.67f * 243.5f)) / (d * d);
        R den = p - R(0.378f) * es;
        return ((R(0.622f) * p) * desdT) / (den * den);
    }
    R rh_interp(R z) const {
        int n = rh_n;
        if (n <= 0) return R(0);
        if (z <= rh_z[0]) return rh_v[0];
        if (z >= rh_z[n - 1]) return rh_v[n - 1];
        int k = 0;
        while (z >= rh_z[k + 1]) ++k;
        R w = (z - rh_z[k]) / (rh_z[k + 1] - rh_z[k]);
        return rh_v[k] + (rh_v[k + 1] - rh_v[k]) * w;
    }
    // Background qv; ASPH_QV_REF_EXP restates asph_h100.cu:149-153.
    R qv_ref(R z) const {
        if (qv_mode == ASPH_QV_REF_EXP) {
            R q = qv0 * mexp<R>(-z / qv_hs);
            R d = mmax<R>(z - qv_zt, R(0));
            return q * mexp<R>(-d / qv_ht);
        }
        if (qv_mode == ASPH_QV_RH_TABLE) {
            R ex = exner_ref(z);
            R T = theta_ref(z) * ex;
            R p = P0 * mpow<R>(ex, cp_rd);
            return rh_interp(z) * qsat(T, p);
        }
        return R(0);
    }
    // --- spline, asph_h100.cu:125-134 --------------------------------------------------------
    // canonical: density uses f(q) = W/sigma (sigma cancels in rho = S1/S2, :331), Horner form.
    R fker(R q) const {
        if (q >= R(2)) return R(0);
        if (q < R(1)) return mfma<R>(q * q, mfma<R>(R(0.75f), q, R(-1.5f)), R(1));
        R t = R(2) - q;
        return (R(0.25f) * t) * (t * t);
    }
    // canonical: k(q) = dW/dq / q / h^5 (:462 generalised to h != 1; equal at h = 1).
    R kgrad(R q) const {
        if (q >= R(2)) return R(0);
        if (q < R(1)) return mfma<R>(R(2.25f), q, R(-3.0f)) * kin;
        R t = R(2) - q;
        return ((kout * t) * t) / q;
    }
    // Clamped cell coordinates by truncation (keys, stencil centre: :236-238, :297-299).
    void cell_trunc(const R* p, int& ix, int& iy, int& iz) const {
        ix = std::min(std::max((int)(p[0] * inv_dx), 0), nx - 1);
        iy = std::min(std::max((int)(p[1] * inv_dy), 0), ny - 1);
        iz = std::min(std::max((int)(p[2] * inv_dz), 0), nz - 1);
    }
    // cell_of by floor, asph_h100.cu:193-198.
    int64_t cell_of(const R* p) const {
        int ix = std::min(std::max((int)std::floor(p[0] * inv_dx), 0), nx - 1);
        int iy = std::min(std::max((int)std::floor(p[1] * inv_dy), 0), ny - 1);
        int iz = std::min(std::max((int)std::floor(p[2] * inv_dz), 0), nz - 1);
        return ((int64_t)iz * ny + iy) * nx + ix;
    }
    static int wrap(int i, int n) { i %= n; return i < 0 ? i + n : i; }   // :199-200
};

template <class R> struct Oracle {
    asph_config cfg;
    Params<R> P;
    int64_t N, ncell, ncol;
    int end_bit;
    std::vector<R> pos, vel, thq, acc, rho, rho_alt, prs, dthdt;
    std::vector<int64_t> rain, dep, precip;
    int64_t clipped = 0;
    std::vector<uint32_t> keys, keys_sorted, perm;
    std::vector<int32_t> cell_range, surf_list;
    float umax[3] = {0, 0, 0};
    // clock (asph_h100.cu:937-942)
    double simt = 0.0, since_sort = 1e9;
    float sort_dt = 0.5f;
    bool initialized = false;
    // tape hooks of the FP64 host adjoint (null otherwise): density sums (S1, S2) and the viscous
    // acceleration sums times mm (x, y, z, 0) of the last evaluation
    std::vector<R>* rec_s12 = nullptr;
    std::vector<R>* rec_av = nullptr;

    explicit Oracle(const asph_config& c) : cfg(c), P(c) {
        N = P.N;
        ncell = N;
        ncol = (int64_t)P.nx * P.ny;
        end_bit = sort_end_bit(P.nx, P.ny, P.nz);
        pos.assign(4 * N, R(0)); vel.assign(4 * N, R(0)); thq.assign(4 * N, R(0));
        acc.assign(4 * N, R(0));
        rho.assign(N, R(0)); rho_alt.assign(N, R(0)); prs.assign(N, R(0)); dthdt.assign(N, R(0));
        rain.assign(ncell, 0); dep.assign(ncell, 0); precip.assign(ncol, 0);
        keys.assign(N, 0); keys_sorted.assign(N, 0); perm.assign(N, 0);
        cell_range.assign(2 * ncell, 0);
    }

    bool flag(uint32_t f) const { return (P.flags & f) != 0; }

    // Fixed-point unit of the rain grid: 2^-40 particle masses in the FP32 instance (as the device,
    // for bit parity); 2^-56 in the FP64 instance, whose rain finite differences would otherwise
    // be limited by the rounding of each deposit to a unit.
    static constexpr double FXS = std::is_same<R, float>::value ? FX_SCALE : 72057594037927936.0;
    static int64_t fx(double v) { return (int64_t)std::llrint(v * FXS); }

    // ---- neighbour structure: asph_h100.cu:229-278, 962-976 -------------------------------
    void sort() {
        for (int64_t i = 0; i < N; ++i) {          // kernel_morton_keys :229-240
            int ix, iy, iz;
            P.cell_trunc(&pos[4 * i], ix, iy, iz);
            keys[i] = morton3((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
        }
        // SortPairs over bits [0,end_bit) of (key, iota): stable argsort (:869, :966)
        std::vector<uint32_t> idx(N);
        std::iota(idx.begin(), idx.end(), 0u);
        uint32_t mask = end_bit >= 32 ? 0xffffffffu : ((1u << end_bit) - 1u);
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
            return (keys[a] & mask) < (keys[b] & mask);
        });
        for (int64_t i = 0; i < N; ++i) { perm[i] = idx[i]; keys_sorted[i] = keys[idx[i]]; }
        // kernel_gather :245-256 (+ acc with ASPH_F_GATHER_ACC)
        auto gather4 = [&](std::vector<R>& a) {
            std::vector<R> b(a.size());
            for (int64_t i = 0; i < N; ++i)
                for (int c = 0; c < 4; ++c) b[4 * i + c] = a[4 * (int64_t)perm[i] + c];
            a.swap(b);
        };
        gather4(pos); gather4(vel); gather4(thq);
        if (flag(ASPH_F_GATHER_ACC)) gather4(acc);
        {
            std::vector<R> b(N);
            for (int64_t i = 0; i < N; ++i) b[i] = rho[perm[i]];
            rho.swap(b);
        }
        // memset + kernel_build_pairs :257-278
        std::fill(cell_range.begin(), cell_range.end(), 0);
        surf_list.clear();
        int64_t surf_cap = 2 * ncol;
        for (int64_t i = 0; i < N; ++i) {
            int64_t cid = P.cell_of(&pos[4 * i]);
            if (i == 0) cell_range[2 * cid] = 0;
            else {
                int64_t pcid = P.cell_of(&pos[4 * (i - 1)]);
                if (cid != pcid) { cell_range[2 * cid] = (int32_t)i; cell_range[2 * pcid + 1] = (int32_t)i; }
            }
            if (i == N - 1) cell_range[2 * cid + 1] = (int32_t)N;
            if ((int)(pos[4 * i + 2] * P.inv_dz) == 0 && (int64_t)surf_list.size() < surf_cap)
                surf_list.push_back((int32_t)i);
        }
    }

    // ---- KDK pieces: asph_h100.cu:535-571, 518-530 ------------------------------------------
    void kick(R hdt) {
        for (int64_t i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c)
                vel[4 * i + c] = mfma<R>(acc[4 * i + c], hdt, vel[4 * i + c]);  // canonical :542
    }
    void drift(R dt) {
        for (int64_t i = 0; i < N; ++i) {
            R* p = &pos[4 * i];
            R* v = &vel[4 * i];
            p[0] = mfma<R>(v[0], dt, p[0]);                                      // canonical :552
            p[1] = mfma<R>(v[1], dt, p[1]);
            p[2] = mfma<R>(v[2], dt, p[2]);
            // :554-557, bounded at 8 passes (identical for |v dt| < 8 L; stops on inf)
            for (int k = 0; k < 8 && p[0] < R(0); ++k) p[0] += P.Lx;
            for (int k = 0; k < 8 && p[0] >= P.Lx; ++k) p[0] -= P.Lx;
            for (int k = 0; k < 8 && p[1] < R(0); ++k) p[1] += P.Ly;
            for (int k = 0; k < 8 && p[1] >= P.Ly; ++k) p[1] -= P.Ly;
            if (p[2] < R(0)) { p[2] = -p[2]; v[2] = std::fabs(v[2]); }            // :558
            if (p[2] > P.Lz) { p[2] = R(2) * P.Lz - p[2]; v[2] = -std::fabs(v[2]); }  // :559
        }
    }
    void theta_update(R dt) {
        for (int64_t i = 0; i < N; ++i)                                          // :569, canonical fma
            thq[4 * i] = mmin<R>(mmax<R>(mfma<R>(dthdt[i], dt, thq[4 * i]), R(100)), P.theta_cap);
    }
    void sponge(R dt) {
        R zs = (R(1) - P.sponge_frac) * P.Lz;                                   // :525
        for (int64_t i = 0; i < N; ++i) {
            R z = pos[4 * i + 2];
            if (z < zs) continue;
            R t = (z - zs) / (P.sponge_frac * P.Lz);
            R sg = (P.sponge_sigma * t) * t;
            vel[4 * i + 2] *= R(1) / (R(1) + sg * dt);                           // :529
        }
    }

    // ---- surface flux: asph_h100.cu:606-626, host side :989-992 ------------------------------
    void surface_flux_params(double t, R& Ts, R& qs_sfc) const {
        R r = mmin<R>(mmax<R>(R((float)t) / P.ramp, R(0)), R(1));
        R sw = (r * r) * (R(3) - R(2) * r);
        Ts = P.th0 + P.dts_skin * sw;
        qs_sfc = P.qsat(Ts, P.pres_ref(R(0)));
    }
    void surface_flux(R dt, double t) {
        R Ts, qs;
        surface_flux_params(t, Ts, qs);
        for (int32_t i : surf_list) {
            R* p = &pos[4 * (int64_t)i];
            if (p[2] >= P.dz) continue;
            R* tq = &thq[4 * (int64_t)i];
            R Pi = mpow<R>(mmax<R>(P.pres_ref(p[2]), R(1)) / P.P0, P.rd_cp);
            R Ta = tq[0] * Pi;
            tq[0] = tq[0] + (((P.c_h * P.v_bg) * (Ts - Ta)) * dt) / P.dz;
            tq[1] = tq[1] + ((((P.beta_moist * P.c_e) * P.v_bg) * (qs - tq[1])) * dt) / P.dz;
            tq[0] = mmin<R>(mmax<R>(tq[0], R(100)), P.theta_cap);
            tq[1] = mmin<R>(mmax<R>(tq[1], R(0)), P.qv_cap);
        }
    }

    // Stencil bounds around particle p (asph_h100.cu:297-303).
    void stencil(const R* p, int& ix, int& iy, int& iz, int& z0, int& z1) const {
        P.cell_trunc(p, ix, iy, iz);
        int S = P.S;
        bool interior = (iz >= S && iz < P.nz - S);
        z0 = interior ? iz - S : std::max(iz - S, 0);
        z1 = interior ? iz + S : std::min(iz + S, P.nz - 1);
    }
    // Separation x_i - x_j in cells (asph_h100.cu:315-318, 453-456 scaled by 1/dx, 1/dy, 1/dz).
    // canonical: each coordinate is scaled to cells first, (x_i / dx) - (x_j / dx) for
    // (x_i - x_j) / dx, and the minimum image (off in the reference) shifts by the period in
    // cells, nx for Lx.
    void csep(const R* pi, const R* pj, R& sx, R& sy, R& sz) const {
        sx = pi[0] * P.inv_dx - pj[0] * P.inv_dx;
        sy = pi[1] * P.inv_dy - pj[1] * P.inv_dy;
        sz = pi[2] * P.inv_dz - pj[2] * P.inv_dz;
        if (flag(ASPH_F_MIN_IMAGE)) {
            if (sx > P.hnx) sx -= P.fnx; else if (sx < -P.hnx) sx += P.fnx;
            if (sy > P.hny) sy -= P.fny; else if (sy < -P.hny) sy += P.fny;
        }
    }

    // ---- density: asph_h100.cu:286-333 ------------------------------------------------------
    void density() {
        const int S = P.S;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            const R* pi = &pos[4 * i];
            int ix, iy, iz, z0, z1;
            stencil(pi, ix, iy, iz, z0, z1);
            R S1 = 0, S2 = 0;
            for (int jz = z0; jz <= z1; ++jz)
                for (int jy = iy - S; jy <= iy + S; ++jy) {
                    int cy = Params<R>::wrap(jy, P.ny);
                    int64_t row = ((int64_t)jz * P.ny + cy) * P.nx;
                    for (int jx = ix - S; jx <= ix + S; ++jx) {
                        int64_t c = row + Params<R>::wrap(jx, P.nx);
                        for (int32_t j = cell_range[2 * c]; j < cell_range[2 * c + 1]; ++j) {
                            const R* pj = &pos[4 * (int64_t)j];
                            R dxm, dym, dzm;
                            csep(pi, pj, dxm, dym, dzm);
                            R r2 = mfma<R>(dzm, dzm, mfma<R>(dym, dym, dxm * dxm));
                            if (r2 < P.lim) {
                                R W = P.fker(msqrt<R>(r2) * P.inv_h);
                                S1 += W;
                                // canonical: W * (1/max(rho_j,1e-3)) for W / max(rho_j,1e-3) (:322)
                                S2 += W * (R(1) / mmax<R>(rho[j], R(1e-3f)));
                            }
                        }
                    }
                }
            if (rec_s12) { (*rec_s12)[2 * i] = S1; (*rec_s12)[2 * i + 1] = S2; }
            R rref = P.rho_ref(pi[2]);
            R rr = S2 > R(1e-20f) ? S1 / S2 : rref;                                    // :331
            rho_alt[i] = mmin<R>(mmax<R>(rr, R(0.5f) * rref), R(2.0f) * rref);   // :332
        }
        rho.swap(rho_alt);   // :1003 swap: reads the old rho, rho becomes the new density
    }

    // ---- EOS + microphysics: asph_h100.cu:340-402 (+ Kessler accretion / evaporation) -------
    void eos_micro(R dt) {
        const R auto_f = R(1) - mexp<R>(-P.rate_auto * dt);   // (1 - exp(-k dt)) of :390
        const R LvCp = P.L_V / P.CP;
        const bool do_auto = flag(ASPH_F_AUTOCONV), do_acc = flag(ASPH_F_ACCRETION),
                   do_evap = flag(ASPH_F_RAIN_EVAP);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            const R* p = &pos[4 * i];
            R* t = &thq[4 * i];
            R z = p[2];
            R r = mmax<R>(rho[i], R(1e-3f));
            R rref = P.rho_ref(z);
            R dd = r - rref;
            dd = mmin<R>(mmax<R>(dd, R(-0.3f) * rref), R(0.3f) * rref);         // :359
            R ppr = P.cs2 * dd;                                                  // :360
            R pfl = mmax<R>(P.pres_ref(z) + ppr, R(1));                          // :361
            R Pi = mpow<R>(pfl / P.P0, P.rd_cp);                                 // :362
            R th = t[0], qv = t[1], qc = t[2];
            R T = th * Pi;
            R qs = P.qsat(T, pfl);
            if (qv > qs || qc > R(0)) {                                          // :370-385
                R dqc = 0;
                for (int it = 0; it < 3; ++it) {
                    R Tn = T + LvCp * dqc;
                    R esn = P.esat(Tn);
                    R qsn = P.qsat(Tn, pfl);
                    R f = (qv - dqc) - qsn;
                    R fp = R(-1) - P.dqsdT(Tn, esn, pfl) * LvCp;
                    dqc -= f / fp;
                    dqc = mmin<R>(dqc, qv);
                    dqc = mmax<R>(dqc, -qc);
                }
                qv -= dqc; qc += dqc;
                th += (P.L_V * dqc) / (P.CP * Pi);
                T += LvCp * dqc;
            }
            int64_t cid = -1;
            int64_t dfx = 0;
            if (do_auto && qc > P.qc_auto) {                                     // :389-394
                R dqr = (qc - P.qc_auto) * auto_f;
                qc -= dqr;
                dfx += fx((double)dqr);
            }
            if (do_acc || do_evap) {
                cid = P.cell_of(p);
                // the FP32 instance reads the grid as the device does (rounded to float); the FP64
                // instance, the finite-difference reference of the adjoint, reads it exactly
                R rf = std::is_same<R, float>::value ? R((float)((double)rain[cid] / FXS))
                                                     : R((double)rain[cid] / FXS);
                R qr = (rf * P.mm) / r;
                if (do_acc && qc > R(0) && qr > R(0)) {
                    // Kessler accretion k2 qc qr^0.875, implicit in qc
                    R qcn = qc / (R(1) + (P.k_accr * dt) * mpow<R>(qr, R(0.875f)));
                    dfx += fx((double)(qc - qcn));
                    qc = qcn;
                }
                if (do_evap && qr > R(0)) {
                    // Kessler rain evaporation (Klemp & Wilhelmson 1978, eq. 2.14; rho in g/cm^3)
                    R qs2 = P.qsat(T, pfl);
                    if (qv < qs2) {
                        R rc = R(0.001f) * r;
                        R rq = rc * qr;
                        R C = R(1.6f) + R(124.9f) * mpow<R>(rq, R(0.2046f));
                        R E = (((R(1) - qv / qs2) * C) * mpow<R>(rq, R(0.525f))) /
                              ((R(5.4e5f) + R(2.55e8f) / (pfl * qs2)) * rc);
                        R dqe = mmin<R>(mmin<R>(E * dt, qr), qs2 - qv);
                        qv += dqe;
                        th -= (P.L_V * dqe) / (P.CP * Pi);
                        dfx -= fx((double)dqe);
                    }
                }
            }
            if (dfx != 0) {
                if (cid < 0) cid = P.cell_of(p);
#pragma omp atomic
                dep[cid] += dfx;
            }
            th = mmin<R>(mmax<R>(th, R(100)), P.theta_cap);                       // :396-398
            qv = mmin<R>(mmax<R>(qv, R(0)), P.qv_cap);
            qc = mmax<R>(qc, R(0));
            t[0] = th; t[1] = qv; t[2] = qc; t[3] = R(0);                         // :400
            prs[i] = ppr;                                                         // :401
        }
    }

    // ---- momentum: asph_h100.cu:415-511 -----------------------------------------------------
    void momentum() {
        const int S = P.S;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            const R* pi = &pos[4 * i];
            const R* vi = &vel[4 * i];
            const R* ti = &thq[4 * i];
            int ix, iy, iz, z0, z1;
            stencil(pi, ix, iy, iz, z0, z1);
            R rhoi = mmax<R>(rho[i], R(1e-3f));
            R inv_rhoi = R(1) / rhoi;
            R pri = prs[i];
            R ax = 0, ay = 0, az = 0, avx = 0, avy = 0, avz = 0;
            for (int jz = z0; jz <= z1; ++jz)
                for (int jy = iy - S; jy <= iy + S; ++jy) {
                    int cy = Params<R>::wrap(jy, P.ny);
                    int64_t row = ((int64_t)jz * P.ny + cy) * P.nx;
                    for (int jx = ix - S; jx <= ix + S; ++jx) {
                        int64_t c = row + Params<R>::wrap(jx, P.nx);
                        for (int32_t j = cell_range[2 * c]; j < cell_range[2 * c + 1]; ++j) {
                            if (j == i) continue;
                            const R* pj = &pos[4 * (int64_t)j];
                            R dxt, dyt, dzt;
                            csep(pi, pj, dxt, dyt, dzt);
                            R r2t = mfma<R>(dzt, dzt, mfma<R>(dyt, dyt, dxt * dxt));
                            if (r2t >= P.lim || r2t < R(1e-12f)) continue;            // :458
                            R q = msqrt<R>(r2t) * P.inv_h;
                            // canonical: physical separation (x_i - x_j) as cells times dx
                            R rx = dxt * P.dx, ry = dyt * P.dy, rz = dzt * P.dz;
                            R rhoj = mmax<R>(rho[j], R(1e-3f));
                            R inv_rhoj = R(1) / rhoj;
                            R prj = prs[j];
                            // canonical: (prj-pri)*(1/rhoi * 1/rhoj) for (prj-pri)/(rhoi*rhoj) (:473)
                            R dp = (prj - pri) * (inv_rhoi * inv_rhoj);
                            const R* vj = &vel[4 * (int64_t)j];
                            R vdx = vi[0] - vj[0], vdy = vi[1] - vj[1], vdz = vi[2] - vj[2];
                            R vdotr = mfma<R>(vdz, rz, mfma<R>(vdy, ry, vdx * rx));
                            // canonical: the viscous denominator with rp*rp -> rp2 (:481-486) and
                            // one division X = 1/(q D) for both quotients: dW/dq/q (:462, outer
                            // piece) = k_out t^2 (D X) and mu/(rbar^2 rhoj) = c vdotr (q X)
                            R rp2 = mfma<R>(rz, rz, mfma<R>(ry, ry, rx * rx));
                            R rbar = R(0.5f) * (rhoi + rhoj);
                            R D = ((rp2 + P.eta2) * (rbar * rbar)) * rhoj;
                            R X = R(1) / (q * D);
                            R t = R(2) - q;
                            R k = q >= R(2) ? R(0)
                                : (q < R(1) ? P.kgrad(q) : ((P.kout * t) * t) * (D * X));
                            R visc = 0;
                            if (vdotr < R(0)) visc = (P.visc_c * vdotr) * (q * X);        // :480-487
                            R br = dp + visc;
                            // canonical: the gradient's 1/dx (:463-465) leaves the sum for the
                            // factor gmx = mm/dx (:497): sum of -(br k) dxt, times gmx
                            R tb = br * k, tv = visc * k;
                            ax = mfma<R>(-tb, dxt, ax); ay = mfma<R>(-tb, dyt, ay); az = mfma<R>(-tb, dzt, az);
                            avx = mfma<R>(-tv, dxt, avx); avy = mfma<R>(-tv, dyt, avy);
                            avz = mfma<R>(-tv, dzt, avz);
                        }
                    }
                }
            ax *= P.gmx; ay *= P.gmy; az *= P.gmz;                                     // :497-499
            avx *= P.gmx; avy *= P.gmy; avz *= P.gmz;
            if (rec_av) {
                (*rec_av)[4 * i] = avx; (*rec_av)[4 * i + 1] = avy; (*rec_av)[4 * i + 2] = avz;
                (*rec_av)[4 * i + 3] = R(0);
            }
            R thv = ti[0] * (R(1) + R(0.608f) * ti[1]);                                // :502-504
            R thvr = P.theta_ref(pi[2]) * (R(1) + R(0.608f) * P.qv_ref(pi[2]));
            az = az + P.GRAV * ((thv - thvr) / thvr - ti[2]);
            R* a = &acc[4 * i];
            a[0] = ax; a[1] = ay; a[2] = az; a[3] = R(0);
            R Pi = mpow<R>(mmax<R>(P.pres_ref(pi[2]) + pri, R(1)) / P.P0, P.rd_cp);    // :508
            R dth = -mfma<R>(vi[2], avz, mfma<R>(vi[1], avy, vi[0] * avx)) / (P.CP * Pi);
            dthdt[i] = mmin<R>(mmax<R>(dth, -P.dthdt_cap), P.dthdt_cap);              // :510
        }
    }

    // ---- rain sedimentation: asph_h100.cu:577-599, exact fixed point ------------------------
    void rain_merge() {
        for (int64_t c = 0; c < ncell; ++c) { rain[c] += dep[c]; dep[c] = 0; }
    }
    int64_t rain_cur(int64_t c) {
        int64_t cur = rain[c] + dep[c];
        return cur < 0 ? 0 : cur;
    }
    void rain_adv(R dt) {
        if (!flag(ASPH_F_SEDIMENT)) {
            for (int64_t c = 0; c < ncell; ++c) {
                int64_t cur = rain[c] + dep[c];
                if (cur < 0) { clipped += -cur; cur = 0; }
                rain[c] = cur;
                dep[c] = 0;
            }
            return;
        }
        float frac = (float)(((float)P.v_rain * (float)dt) * (float)P.inv_dz);   // :587-588
        if (frac > 1.f) frac = 1.f;
        std::vector<int64_t> out(ncell);
        const int64_t plane = ncol;
        for (int64_t c = 0; c < ncell; ++c) {
            int64_t cur = rain[c] + dep[c];
            if (cur < 0) { clipped += -cur; cur = 0; }
            int64_t om = (int64_t)((double)cur * (double)frac);
            int64_t im = 0;
            if (c + plane < ncell) im = (int64_t)((double)rain_cur(c + plane) * (double)frac);
            out[c] = cur - om + im;
            if (c < plane) precip[c] += om;
        }
        rain.swap(out);
        std::fill(dep.begin(), dep.end(), 0);
    }

    // ---- reductions: asph_h100.cu:631-675 ---------------------------------------------------
    void max_vel(float out[3]) const {
        R m[3] = {0, 0, 0};
        for (int64_t i = 0; i < N; ++i)
            for (int c = 0; c < 3; ++c) m[c] = mmax<R>(m[c], std::fabs(vel[4 * i + c]));
        for (int c = 0; c < 3; ++c) out[c] = (float)m[c];
    }
    void clamp_vel(R vmax) {
        for (int64_t i = 0; i < N; ++i) {
            R* v = &vel[4 * i];
            R s2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (s2 > vmax * vmax) {
                R s = msqrt<R>(s2), f = vmax / s;
                v[0] *= f; v[1] *= f; v[2] *= f;
            }
        }
    }
    void audit(double out[7]) const {
        double ke = 0, pe = 0, ie = 0, w = 0, pz = 0;
        for (int64_t i = 0; i < N; ++i) {
            const R* p = &pos[4 * i];
            const R* v = &vel[4 * i];
            const R* t = &thq[4 * i];
            R Pi = mpow<R>(mmax<R>(P.pres_ref(p[2]), R(1)) / P.P0, P.rd_cp);
            ke += 0.5 * (double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            pe += (double)(P.GRAV * p[2]);
            ie += (double)((P.CP * t[0]) * Pi);
            w += (double)(t[1] + t[2]);
            pz += (double)v[2];
        }
        double m = (double)P.pmass;
        out[0] = ke * m; out[1] = pe * m; out[2] = ie * m; out[3] = w * m; out[4] = pz * m;
        out[5] = m * (double)N;
        double rw = 0;
        for (int64_t c = 0; c < ncell; ++c) rw += (double)(rain[c] + dep[c]);
        for (int64_t c = 0; c < ncol; ++c) rw += (double)precip[c];
        out[6] = rw / FXS * m;
    }

    // ---- dumps: asph_h100.cu:680-723 --------------------------------------------------------
    void coarse_dump(int sx, int sy, int sz, std::vector<float>& out) const {
        int cx = (P.nx + sx - 1) / sx, cy = (P.ny + sy - 1) / sy, cz = (P.nz + sz - 1) / sz;
        out.assign((size_t)cx * cy * cz * 6, 0.f);
        for (int idx = 0; idx < cx * cy * cz; ++idx) {
            int cxx = idx % cx, cyy = (idx / cx) % cy, czz = idx / (cx * cy);
            float a[6] = {0, 0, 0, 0, 0, 0};
            int n = 0;
            int ix1 = std::min((cxx + 1) * sx, P.nx), iy1 = std::min((cyy + 1) * sy, P.ny),
                iz1 = std::min((czz + 1) * sz, P.nz);
            for (int iz = czz * sz; iz < iz1; ++iz)
                for (int iy = cyy * sy; iy < iy1; ++iy)
                    for (int ix = cxx * sx; ix < ix1; ++ix) {
                        int64_t c = ((int64_t)iz * P.ny + iy) * P.nx + ix;
                        for (int32_t j = cell_range[2 * c]; j < cell_range[2 * c + 1]; ++j) {
                            a[0] += (float)thq[4 * (int64_t)j]; a[1] += (float)thq[4 * (int64_t)j + 1];
                            a[2] += (float)thq[4 * (int64_t)j + 2];
                            a[3] += (float)vel[4 * (int64_t)j]; a[4] += (float)vel[4 * (int64_t)j + 1];
                            a[5] += (float)vel[4 * (int64_t)j + 2];
                            n++;
                        }
                    }
            float inv = n ? 1.f / (float)n : 0.f;
            for (int c = 0; c < 6; ++c) out[(size_t)idx * 6 + c] = a[c] * inv;
        }
    }
    void column_diag(std::vector<float>& out) const {
        out.assign((size_t)ncol * 4, 0.f);
        for (int64_t col = 0; col < ncol; ++col) {
            int iy = (int)(col / P.nx), ix = (int)(col % P.nx);
            float th = 0, qv = 0, qlmax = 0, wmax = 0;
            int n = 0;
            for (int k = 0; k < P.nz; ++k) {
                int64_t c = ((int64_t)k * P.ny + iy) * P.nx + ix;
                for (int32_t j = cell_range[2 * c]; j < cell_range[2 * c + 1]; ++j) {
                    th += (float)thq[4 * (int64_t)j]; qv += (float)thq[4 * (int64_t)j + 1];
                    qlmax = std::fmax(qlmax, (float)thq[4 * (int64_t)j + 2]);
                    wmax = std::fmax(wmax, std::fabs((float)vel[4 * (int64_t)j + 2]));
                    n++;
                }
            }
            if (n) { float inv = 1.f / (float)n; th *= inv; qv *= inv; }
            out[4 * col] = th; out[4 * col + 1] = qv; out[4 * col + 2] = qlmax; out[4 * col + 3] = wmax;
        }
    }

    // ---- orchestration: asph_h100.cu:913-920 and 962-1038 -----------------------------------
    void force(R dt) {
        density();
        eos_micro(dt);
        rain_merge();   // the init-time autoconversion lands in the rain grid (:918)
        momentum();
        initialized = true;
    }
    void step(R dt, double t, bool do_sort) {
        if (do_sort) sort();
        std::fill(dep.begin(), dep.end(), 0);
        kick(R(0.5f) * dt);
        drift(dt);
        if (flag(ASPH_F_SURFACE_FLUX)) surface_flux(dt, t);
        density();
        eos_micro(dt);
        momentum();
        theta_update(dt);
        kick(R(0.5f) * dt);
        if (flag(ASPH_F_SPONGE)) sponge(dt);
        rain_adv(dt);
        max_vel(umax);
    }
};

}  // namespace oracle
