// asph_internal.h - parameters and physics helpers shared by the host library and the gfx950
// kernels. All arithmetic here is the project's canonical form of the reference formulas
// (asph_h100.cu:125-200): fma only where written, exp/log/pow from asph_fmath.h; the file is
// compiled with -ffp-contract=off on both host and device so both evaluate it identically.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "asph.h"
#include "asph_fmath.h"

#define ASPH_DH __host__ __device__ __forceinline__

namespace asph {

// Physical constants, asph_h100.cu:56, 66-70.
constexpr float GRAV = 9.80665f;
constexpr float R_D = 287.05f;
constexpr float CP = 1004.0f;
constexpr float P0 = 100000.0f;
constexpr float L_V = 2.501e6f;
constexpr float SIGMA = 0.31830988618f;
// Rain is held as int64 in units of one particle mass times 2^-40 (exact sums, exact
// sedimentation mass balance, order-independent deposits).
constexpr double FX_SCALE = 1099511627776.0;
constexpr double FX_INV = 1.0 / 1099511627776.0;

// Derived constants of a configuration (evaluated once on the host in float).
struct DevParams {
    int32_t nx, ny, nz, S;
    int32_t N, ncell, ncol, end_bit;
    uint32_t flags;
    float dx, dy, dz, inv_dx, inv_dy, inv_dz, Lx, Ly, Lz, hLx, hLy;
    float fnx, fny, hnx, hny;   // periods in x and y [cells] and their halves
    float inv_h, lim, kin, kout;
    float gmx, gmy, gmz;        // mm / dx, mm / dy, mm / dz: the momentum sums' factors
    float vcell, rho0, pmass, mm, eta2, visc_c, cs, cs2;
    float th0, n2, exA, cp_rd, rd_cp;
    float sponge_sigma, sponge_frac, sponge_zs, sponge_depth, dthdt_cap;
    float c_h, c_e, beta_moist, v_bg, dts_skin, ramp;
    float qc_auto, rate_auto, k_accr, v_rain, theta_cap, qv_cap;
    int32_t qv_mode;
    float qv0, qv_hs, qv_zt, qv_ht;
    int32_t rh_n;
    float rh_z[ASPH_RH_KNOTS], rh_v[ASPH_RH_KNOTS];
};

// --- reference state, asph_h100.cu:136-153 ------------------------------------------------------
ASPH_DH float theta_ref(const DevParams& P, float z) {
    if (P.n2 > 0.0f) return P.th0 * asphm::fexp((P.n2 * z) / GRAV);
    return P.th0;
}
ASPH_DH float exner_ref(const DevParams& P, float z) {
    if (P.n2 > 0.0f) return 1.0f - P.exA * (1.0f - asphm::fexp((-P.n2 * z) / GRAV));
    return 1.0f - (GRAV * z) / (CP * P.th0);
}
ASPH_DH float pres_of_exner(const DevParams& P, float ex) { return P0 * asphm::fpow(ex, P.cp_rd); }
ASPH_DH float pres_ref(const DevParams& P, float z) { return pres_of_exner(P, exner_ref(P, z)); }
ASPH_DH float rho_ref(const DevParams& P, float z) {
    return pres_ref(P, z) / ((R_D * theta_ref(P, z)) * exner_ref(P, z));
}
// --- thermodynamics, asph_h100.cu:155-171 ---------------------------------------------------------
ASPH_DH float esat(float T) {
    float Tc = T - 273.15f;
    return 611.2f * asphm::fexp((17.67f * Tc) / (Tc + 243.5f));
}
ASPH_DH float qsat(float T, float p) {
    float es = fminf(esat(T), 0.5f * p);
    float den = p - 0.378f * es;
    return den > 1.0f ? (0.622f * es) / den : 0.0f;
}
ASPH_DH float dqsdT(float T, float es, float p) {
    float Tc = T - 273.15f;
    float d = Tc + 243.5f;
    float desdT = (es * (17.67f * 243.5f)) / (d * d);
    float den = p - 0.378f * es;
    return ((0.622f * p) * desdT) / (den * den);
}
// Piecewise-linear RH(z) over non-decreasing knots, clamped outside. The segment is the last
// k <= n-2 with z >= rh_z[k] (where the scan k = 0, 1, ... while z >= rh_z[k+1] stops), picked
// with constant indices only so the table can live in registers.
ASPH_DH float rh_interp(const DevParams& P, float z) {
    const int n = P.rh_n;
    if (n <= 0) return 0.0f;
    if (z <= P.rh_z[0]) return P.rh_v[0];
    float zl = P.rh_z[0], vl = P.rh_v[0];
#pragma unroll
    for (int k = 1; k < ASPH_RH_KNOTS; ++k) {
        if (k == n - 1) { zl = P.rh_z[k]; vl = P.rh_v[k]; }
    }
    if (z >= zl) return vl;
    float za = P.rh_z[0], va = P.rh_v[0], zb = P.rh_z[1], vb = P.rh_v[1];
#pragma unroll
    for (int k = 1; k + 1 < ASPH_RH_KNOTS; ++k) {
        if (k + 1 < n && z >= P.rh_z[k]) {
            za = P.rh_z[k]; va = P.rh_v[k];
            zb = P.rh_z[k + 1]; vb = P.rh_v[k + 1];
        }
    }
    float w = (z - za) / (zb - za);
    return va + (vb - va) * w;
}
// Background moisture; ASPH_QV_REF_EXP is asph_h100.cu:149-153.
ASPH_DH float qv_ref(const DevParams& P, float z) {
    if (P.qv_mode == ASPH_QV_REF_EXP) {
        float q = P.qv0 * asphm::fexp(-z / P.qv_hs);
        float d = fmaxf(z - P.qv_zt, 0.0f);
        return q * asphm::fexp(-d / P.qv_ht);
    }
    if (P.qv_mode == ASPH_QV_RH_TABLE) {
        float ex = exner_ref(P, z);
        float T = theta_ref(P, z) * ex;
        float p = pres_of_exner(P, ex);
        return rh_interp(P, z) * qsat(T, p);
    }
    return 0.0f;
}

// --- spline, asph_h100.cu:125-134 (canonical forms; sigma dropped in the density ratio) ----------
ASPH_DH float fker(float q) {
    if (q >= 2.0f) return 0.0f;
    if (q < 1.0f) return __builtin_fmaf(q * q, __builtin_fmaf(0.75f, q, -1.5f), 1.0f);
    float t = 2.0f - q;
    return (0.25f * t) * (t * t);
}
ASPH_DH float kgrad(const DevParams& P, float q) {
    if (q >= 2.0f) return 0.0f;
    if (q < 1.0f) return __builtin_fmaf(2.25f, q, -3.0f) * P.kin;
    float t = 2.0f - q;
    return ((P.kout * t) * t) / q;
}
// Branch-free forms for the pair loops: both pieces are evaluated with the operations above and
// the applicable one is selected, so the values equal fker / kgrad. In fker_sel, t = max(2 - q, 0)
// is 2 - q exactly inside the support and 0 beyond it, where the outer piece is +0. kgrad_sel
// keeps the explicit test (the compiler then skips its division where no lane needs it, which
// measured faster than a clamped t). A quotient at q = 0 (a placeholder pair) is never selected.
ASPH_DH float fker_sel(float q) {
    const float inner = __builtin_fmaf(q * q, __builtin_fmaf(0.75f, q, -1.5f), 1.0f);
    const float t = fmaxf(2.0f - q, 0.0f);
    const float outer = (0.25f * t) * (t * t);
    return q < 1.0f ? inner : outer;
}
ASPH_DH float kgrad_sel(const DevParams& P, float q) {
    const float inner = __builtin_fmaf(2.25f, q, -3.0f) * P.kin;
    const float t = 2.0f - q;
    const float outer = ((P.kout * t) * t) / q;
    return q >= 2.0f ? 0.0f : (q < 1.0f ? inner : outer);
}

// --- Morton keys, asph_h100.cu:176-186 ------------------------------------------------------------
ASPH_DH uint32_t part1by2(uint32_t n) {
    n &= 0x000003ffu;
    n = (n | (n << 16)) & 0x030000ffu;
    n = (n | (n << 8)) & 0x0300f00fu;
    n = (n | (n << 4)) & 0x030c30c3u;
    n = (n | (n << 2)) & 0x09249249u;
    return n;
}
ASPH_DH uint32_t compact1by2(uint32_t n) {
    n &= 0x09249249u;
    n = (n | (n >> 2)) & 0x030c30c3u;
    n = (n | (n >> 4)) & 0x0300f00fu;
    n = (n | (n >> 8)) & 0x030000ffu;
    n = (n | (n >> 16)) & 0x000003ffu;
    return n;
}
ASPH_DH uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2);
}
// Clamped truncated cell coordinates (asph_h100.cu:236-238, 297-299).
ASPH_DH void cell_trunc(const DevParams& P, float x, float y, float z, int& ix, int& iy, int& iz) {
    ix = min(max((int)(x * P.inv_dx), 0), P.nx - 1);
    iy = min(max((int)(y * P.inv_dy), 0), P.ny - 1);
    iz = min(max((int)(z * P.inv_dz), 0), P.nz - 1);
}
// cell_of by floor (asph_h100.cu:193-198).
ASPH_DH int cell_of(const DevParams& P, float x, float y, float z) {
    int ix = min(max((int)floorf(x * P.inv_dx), 0), P.nx - 1);
    int iy = min(max((int)floorf(y * P.inv_dy), 0), P.ny - 1);
    int iz = min(max((int)floorf(z * P.inv_dz), 0), P.nz - 1);
    return (iz * P.ny + iy) * P.nx + ix;
}
ASPH_DH int wrapi(int i, int n) { i %= n; return i < 0 ? i + n : i; }

ASPH_DH float rain_fx_to_float(int64_t v) { return (float)((double)v * FX_INV); }
ASPH_DH int64_t to_fx(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __double2ll_rn((double)v * FX_SCALE);
#else
    return (int64_t)__builtin_llrint((double)v * FX_SCALE);
#endif
}

// Host: fill DevParams from a configuration (float evaluation, the order matches the oracle).
void make_params(const asph_config& c, DevParams& P);

// Morton bits of one brick of the neighbour kernels: 8 x 8 x 8 cells, keys [base, base + 512).
constexpr int BRICK_BITS = 9;

// One domain of a handle. A plain run is a handle with one member; an ensemble concatenates its
// members' particles (and cells, one per particle) and columns, and prefixes their Morton keys
// with the member index so that one sort keeps every member contiguous and cell-ordered.
struct MemberDesc {
    DevParams P;
    int32_t p0;     // first particle (= first cell) of the member in the concatenated arrays
    int32_t col0;   // first surface column
    uint32_t k0;    // key offset: member index << key bits
    int32_t first_block;
};

// Work-item map of one launch: the member and local start of a block of 256 particles / cells,
// or of an 8x8x8-cell brick (base = local Morton key base, ord = brick ordinal in the member).
struct BlockMap {
    int32_t m;
    int32_t base;
    int32_t ord;
    int32_t pad;
};

// Surface skin temperature and saturation humidity at simulated time t (asph_h100.cu:989-992).
ASPH_DH void sfc_params(const DevParams& P, double t, float& Ts, float& qs) {
    float r = fminf(fmaxf((float)t / P.ramp, 0.f), 1.f);
    float sw = (r * r) * (3.f - 2.f * r);
    Ts = P.th0 + P.dts_skin * sw;
    qs = qsat(Ts, pres_ref(P, 0.0f));
}
// 1 - exp(-k dt) of the autoconversion (asph_h100.cu:390).
ASPH_DH float auto_factor(const DevParams& P, float dt) { return 1.0f - asphm::fexp(-P.rate_auto * dt); }
// Fraction of a cell's rain leaving through its floor per step (asph_h100.cu:587-588).
ASPH_DH float rain_frac(const DevParams& P, float dt) {
    float f = (P.v_rain * dt) * P.inv_dz;
    return f > 1.f ? 1.f : f;
}

}  // namespace asph
