// asph_nbr.hip - the two neighbour-loop kernels of the step on gfx950: lagged-Shepard density
// fused with the EOS and warm-rain microphysics, and momentum fused with the theta update, the
// second half kick, the sponge and the max|v| reduction.
//
// Work decomposition (mode A): one 512-thread workgroup per 8x8x8-cell brick of a member's
// Morton key space (local keys [base, base+512), listed in the brick map; an ensemble launches
// the bricks of all its members together). The workgroup owns the particles LISTED (by the last
// sort) in its brick's cells: in key order these are one contiguous index range, taken in rounds
// of 512 with one particle per thread, so a brick of any occupancy needs no owned-particle list
// and neighbouring lanes hold neighbouring particles. It stages into LDS every particle listed in
// the brick's halo (brick +- margin cells), stored in row-major order of the halo cells with a
// per-cell start table. The margin is the stencil S when the drift bound d since the sort allows
// (a particle lies within [c - d, c + 1 + d] of its listed cell c, so the culled walk of its
// current cell's stencil stays within c +- floor(1 + 2d + R), R the culling radius), else S + 1.
// For a particle each stencil row (jz, jy) is one contiguous LDS span, and walking rows jz -> jy
// and spans left to right visits candidates exactly in the reference's jz -> jy -> jx -> j order
// (asph_h100.cu:306-327, 444-495): the sums are bit-identical to a thread-per-particle loop over
// the cell table.
//
// A halo holding more than MAXP particles (clumped or compressed regions) is streamed through
// LDS in slabs of consecutive halo rows (z-major, then y): each thread keeps its sums in
// registers and walks, slab by slab, the rows of its stencil that lie in the staged slab, so the
// candidate order is still the reference's. Only a brick with a single halo row beyond MAXP, more
// than NSLAB slabs, or more than 65534 halo particles, runs the thread-per-particle loop over the
// global cell table, as does mode B (a work item whose drift bound reaches 0.95 cell, or a
// domain narrower than a halo).
//
// Candidate culling is exact: particles listed in cell c lie within [c - delta, c + 1 + delta] on
// each axis, delta the work item's drift bound: the largest distance beyond their listed cells of
// the particles listed in the bricks its halo can reach, measured by the kick/drift kernel of the
// step (before its first step after a sort every particle is in its listed cell, and any recorded
// bound holds). Rows whose minimum metric distance exceeds the support, and cells of a row beyond
// the row's chord, cannot hold an in-support particle and are skipped; surviving candidates go
// through the reference test r2 < 4h^2 unchanged.
//
// In the momentum kernel the in-support pair work is separated from the candidate test by a
// per-thread queue of LDS slots (QK entries): the test loop takes candidates two at a time and
// appends, and when a lane of the wavefront could not take two more, all lanes drain theirs in
// order, several entries per iteration with their operands loaded first. Lanes therefore spend the
// expensive pair body only on in-support pairs, in the reference order. The density kernel's pair
// term (a square root and the spline) is cheaper than queueing it: it is evaluated for every
// candidate and added where the support test holds.
#include "asph_adjoint.h"
#include "asph_kernels.h"

namespace asph {

namespace {

constexpr int NB = BRICK_BITS;     // Morton bits per brick: 8 x 8 x 8 cells
constexpr int BS = 8;              // brick edge in cells
constexpr int BT = 512;            // threads per workgroup (one per brick cell)
constexpr int BW = BT / 64;        // wavefronts per workgroup
constexpr int QK = 12;             // queued in-support pairs per thread, momentum
constexpr int MU = 3;              // queue entries per drain iteration, momentum
static_assert(QK % MU == 0 && QK % 2 == 0, "a drain iteration never reads past the queue");
template <bool MOM> constexpr int queue_len() { return MOM ? QK : 0; }   // density queues nothing
constexpr int NSLAB = 64;          // most slabs of one halo (beyond: the global-memory loop)
constexpr int CHUNK_MIN = 256;     // bounds of the work-item size (tuned at run time)
constexpr int CHUNK_MAX = 1024;
constexpr float DELTA_EPS = 1e-3f; // rounding margin of the culling bounds [cells]

// Pair separations are taken in cells: each coordinate is scaled by 1/dx (1/dy, 1/dz) first
// and the scaled coordinates are subtracted, (x_i / dx) - (x_j / dx). The staged halos hold the
// scaled coordinates, so a candidate test needs no multiplication before its squares.
__device__ __forceinline__ float4 cellpos(const DevParams& P, float4 x) {
    return make_float4(x.x * P.inv_dx, x.y * P.inv_dy, x.z * P.inv_dz, x.w);
}

// Minimum image of a separation in cells, `if (s > n/2) s -= n; else if (s < -n/2) s += n;`
// without branches: the shift is selected and added (an added +0 changes nothing but the sign of
// a zero difference, which no sum sees).
__device__ __forceinline__ void min_image(const DevParams& P, float& sx, float& sy) {
    if (P.flags & ASPH_F_MIN_IMAGE) {
        // |s| > n/2: shift by n against the sign of s (the same value as the two tests)
        const float tx = fabsf(sx) > P.hnx ? copysignf(P.fnx, -sx) : 0.f;
        const float ty = fabsf(sy) > P.hny ? copysignf(P.fny, -sy) : 0.f;
        sx += tx;
        sy += ty;
    }
}

// Separation of i and j in cells from their scaled coordinates (minimum image when WRAP).
template <bool WRAP = true>
__device__ __forceinline__ void csep(const DevParams& P, float4 ci, float4 cj, float& sx, float& sy, float& sz) {
    sx = ci.x - cj.x;
    sy = ci.y - cj.y;
    sz = ci.z - cj.z;
    if (WRAP) min_image(P, sx, sy);
}

__device__ __forceinline__ float r2of(float sx, float sy, float sz) {
    return __builtin_fmaf(sz, sz, __builtin_fmaf(sy, sy, sx * sx));
}

// Correctly rounded square root for 0 and arguments above 2^-96 (squared pair distances): the
// hardware estimate (within 1 ulp) corrected by the signs of the residuals x - s*s of its two
// neighbours, the same bits as IEEE sqrtf.
__device__ __forceinline__ float sqrt_rn(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1), sp = __int_as_float(__float_as_int(s) + 1);
    float r = __builtin_fmaf(-sm, s, x) <= 0.f ? sm : s;
    r = __builtin_fmaf(-sp, s, x) > 0.f ? sp : r;
    return r;
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// Stencil z range clipped at the walls, asph_h100.cu:301-303.
template <int S> __device__ __forceinline__ void zrange(const DevParams& P, int iz, int& z0, int& z1) {
    bool interior = (iz >= S && iz < P.nz - S);
    z0 = interior ? iz - S : max(iz - S, 0);
    z1 = interior ? iz + S : min(iz + S, P.nz - 1);
}

// ---------------------------------------------------------------------------------------------
// Thread-per-particle loops over the global cell table (mode B and overflow fallback).
// ---------------------------------------------------------------------------------------------
template <int S>
__device__ __forceinline__ void dens_sum_global(const DevParams& P, uint32_t k0, float4 pi, const float4* __restrict__ pos,
                                const float* __restrict__ rho_old, const uint32_t* __restrict__ cstart,
                                float& S1, float& S2) {
    int ix, iy, iz, z0, z1;
    cell_trunc(P, pi.x, pi.y, pi.z, ix, iy, iz);
    zrange<S>(P, iz, z0, z1);
    const float4 ci = cellpos(P, pi);
    S1 = 0.f;
    S2 = 0.f;
    for (int jz = z0; jz <= z1; ++jz) {
        uint32_t kz = k0 | (part1by2((uint32_t)jz) << 2);
        for (int jy = iy - S; jy <= iy + S; ++jy) {
            uint32_t kyz = kz | (part1by2((uint32_t)wrapi(jy, P.ny)) << 1);
            for (int jx = ix - S; jx <= ix + S; ++jx) {
                uint32_t k = kyz | part1by2((uint32_t)wrapi(jx, P.nx));
                uint32_t s = cstart[k], e = cstart[k + 1];
                for (uint32_t j = s; j < e; ++j) {
                    float dxm, dym, dzm;
                    csep(P, ci, cellpos(P, pos[j]), dxm, dym, dzm);
                    float r2 = r2of(dxm, dym, dzm);
                    if (r2 < P.lim) {
                        float W = fker(sqrt_rn(r2) * P.inv_h);
                        S1 += W;
                        S2 += W * (1.0f / fmaxf(rho_old[j], 1e-3f));
                    }
                }
            }
        }
    }
}

struct MomAcc {
    float ax, ay, az, avx, avy, avz;
};

// One in-support momentum pair (asph_h100.cu:460-491), shared by all paths, from the separation
// in cells (sx, sy, sz) and its square r2t. The physical separation is s times (dx, dy, dz). One
// division X = 1/(q D), D the viscous denominator, gives both quotients: the outer spline piece's
// dW/dq / q = k_out t^2 (D X) and the viscous factor c vdotr / D = c vdotr (q X). The gradient's
// 1/dx stays out of the sums: they take -(br k) s_x, and the epilogue multiplies by mm / dx.
__device__ __forceinline__ void mom_pair(const DevParams& P, float sx, float sy, float sz, float r2t,
                                         float4 vi, float rhoi, float inv_rhoi, float pri, float rhoj,
                                         float inv_rhoj, float4 vj, float prj, MomAcc& A,
                                         bool act = true) {
    const float q = sqrt_rn(r2t) * P.inv_h;
    const float rx = sx * P.dx, ry = sy * P.dy, rz = sz * P.dz;
    const float dp = (prj - pri) * (inv_rhoi * inv_rhoj);
    const float vdx = vi.x - vj.x, vdy = vi.y - vj.y, vdz = vi.z - vj.z;
    const float vdotr = __builtin_fmaf(vdz, rz, __builtin_fmaf(vdy, ry, vdx * rx));
    const float rp2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
    const float rbar = 0.5f * (rhoi + rhoj);
    const float D = ((rp2 + P.eta2) * (rbar * rbar)) * rhoj;   // positive
    const float X = 1.0f / (q * D);
    const float inner = __builtin_fmaf(2.25f, q, -3.0f) * P.kin;
    const float t = 2.0f - q;
    const float outer = ((P.kout * t) * t) * (D * X);
    // q = 0 (a placeholder pair) makes X and the outer piece infinite: never selected
    const float kk = q >= 2.0f ? 0.0f : (q < 1.0f ? inner : outer);
    // evaluated for every pair and kept for approaching ones (vdotr < 0)
    float visc = vdotr < 0.f ? (P.visc_c * vdotr) * (q * X) : 0.f;
    float br = dp + visc;
    // act = false: a placeholder pair of an idle queue lane, evaluated and dropped. Its pressure
    // and viscous factors are replaced by zeros and its kk is finite, so each sum takes
    // fma(-0, s, a) = a for every nonzero a. Only a sum that is exactly zero at that point could
    // change the sign of its zero.
    br = act ? br : 0.f;
    visc = act ? visc : 0.f;
    const float tb = br * kk, tv = visc * kk;
    A.ax = __builtin_fmaf(-tb, sx, A.ax);
    A.ay = __builtin_fmaf(-tb, sy, A.ay);
    A.az = __builtin_fmaf(-tb, sz, A.az);
    A.avx = __builtin_fmaf(-tv, sx, A.avx);
    A.avy = __builtin_fmaf(-tv, sy, A.avy);
    A.avz = __builtin_fmaf(-tv, sz, A.avz);
}

template <int S>
__device__ __forceinline__ void mom_sum_global(const DevParams& P, uint32_t k0, float4 pi, float4 vi, float rhoi, float pri,
                               const float4* __restrict__ pos, const float4* __restrict__ vel,
                               const float* __restrict__ rho, const float* __restrict__ prs,
                               const uint32_t* __restrict__ cstart, MomAcc& A) {
    int ix, iy, iz, z0, z1;
    cell_trunc(P, pi.x, pi.y, pi.z, ix, iy, iz);
    zrange<S>(P, iz, z0, z1);
    const float4 ci = cellpos(P, pi);
    const float inv_rhoi = 1.0f / rhoi;
    for (int jz = z0; jz <= z1; ++jz) {
        uint32_t kz = k0 | (part1by2((uint32_t)jz) << 2);
        for (int jy = iy - S; jy <= iy + S; ++jy) {
            uint32_t kyz = kz | (part1by2((uint32_t)wrapi(jy, P.ny)) << 1);
            for (int jx = ix - S; jx <= ix + S; ++jx) {
                uint32_t k = kyz | part1by2((uint32_t)wrapi(jx, P.nx));
                uint32_t s = cstart[k], e = cstart[k + 1];
                for (uint32_t j = s; j < e; ++j) {
                    float dxt, dyt, dzt;
                    csep(P, ci, cellpos(P, pos[j]), dxt, dyt, dzt);
                    const float r2t = r2of(dxt, dyt, dzt);
                    // j == i has r2t = 0 and is excluded here as by `if (j==i) continue` (:452)
                    if (r2t >= P.lim || r2t < 1e-12f) continue;
                    const float rhoj = fmaxf(rho[j], 1e-3f);
                    mom_pair(P, dxt, dyt, dzt, r2t, vi, rhoi, inv_rhoi, pri, rhoj, 1.0f / rhoj, vel[j],
                             prs[j], A);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Per-particle epilogues.
// ---------------------------------------------------------------------------------------------
// density ratio + clamp (:331-332), surface flux (:606-626), EOS + microphysics (:353-401),
// Kessler accretion and rain evaporation. Writes rho_new, thq, prs and p' into vel.w (the
// momentum kernel reads (v, p') of a neighbour with one load; the reference leaves vel.w unused),
// and (max(rho, 1e-3), 1 / max(rho, 1e-3)) into rinv, the two density factors of a momentum pair
// (asph_h100.cu:455-458), computed once per particle instead of once per pair.
struct DensConst {
    float auto_f, Ts, qs_sfc;
    int p0;          // member's first particle (= first cell)
};

__device__ __forceinline__ void dens_epilogue(const DevParams& P, const DensArgs& A, const DensConst& K, int i,
                              float4 pi, float S1, float S2) {
    const float z = pi.z;
    const float ex = exner_ref(P, z);
    const float pref = pres_of_exner(P, ex);
    const float rref = pref / ((R_D * theta_ref(P, z)) * ex);
    float rr = S2 > 1e-20f ? S1 / S2 : rref;
    float rnew = fminf(fmaxf(rr, 0.5f * rref), 2.0f * rref);
    A.rho_new[i] = rnew;
    {
        const float rc = fmaxf(rnew, 1e-3f);
        A.rinv[i] = make_float2(rc, 1.0f / rc);
    }
    if (A.s12_out) A.s12_out[i] = make_float2(S1, S2);
    if (!(A.bits & 1)) return;

    float4 t = A.thq[i];
    if ((A.bits & 2) && (P.flags & ASPH_F_SURFACE_FLUX) && (A.keys_sorted[i] & 0x24924924u) == 0u &&
        z < P.dz) {
        float Pi = asphm::fpow(fmaxf(pref, 1.0f) / P0, P.rd_cp);
        float Ta = t.x * Pi;
        t.x = t.x + (((P.c_h * P.v_bg) * (K.Ts - Ta)) * A.dt) / P.dz;
        t.y = t.y + ((((P.beta_moist * P.c_e) * P.v_bg) * (K.qs_sfc - t.y)) * A.dt) / P.dz;
        t.x = fminf(fmaxf(t.x, 100.f), P.theta_cap);
        t.y = fminf(fmaxf(t.y, 0.f), P.qv_cap);
    }
    const float r = fmaxf(rnew, 1e-3f);
    float dd = r - rref;
    dd = fminf(fmaxf(dd, -0.3f * rref), 0.3f * rref);
    const float ppr = P.cs2 * dd;
    const float pfl = fmaxf(pref + ppr, 1.0f);
    const float Pi = asphm::fpow(pfl / P0, P.rd_cp);
    float th = t.x, qv = t.y, qc = t.z;
    float T = th * Pi;
    const float LvCp = L_V / CP;
    float qs = qsat(T, pfl);
    if (qv > qs || qc > 0.f) {
        float dqc = 0.f;
        for (int it = 0; it < 3; ++it) {
            float Tn = T + LvCp * dqc;
            float esn = esat(Tn);
            float qsn = qsat(Tn, pfl);
            float f = (qv - dqc) - qsn;
            float fp = -1.f - dqsdT(Tn, esn, pfl) * LvCp;
            dqc -= f / fp;
            dqc = fminf(dqc, qv);
            dqc = fmaxf(dqc, -qc);
        }
        qv -= dqc;
        qc += dqc;
        th += (L_V * dqc) / (CP * Pi);
        T += LvCp * dqc;
    }
    int64_t dfx = 0;
    if ((P.flags & ASPH_F_AUTOCONV) && qc > P.qc_auto) {
        float dqr = (qc - P.qc_auto) * K.auto_f;
        qc -= dqr;
        dfx += to_fx(dqr);
    }
    int cid = -1;
    if (P.flags & (ASPH_F_ACCRETION | ASPH_F_RAIN_EVAP)) {
        cid = K.p0 + cell_of(P, pi.x, pi.y, pi.z);
        float rf = rain_fx_to_float(A.rain[cid]);
        float qr = (rf * P.mm) / r;
        if ((P.flags & ASPH_F_ACCRETION) && qc > 0.f && qr > 0.f) {
            float qcn = qc / (1.0f + (P.k_accr * A.dt) * asphm::fpow(qr, 0.875f));
            dfx += to_fx(qc - qcn);
            qc = qcn;
        }
        if ((P.flags & ASPH_F_RAIN_EVAP) && qr > 0.f) {
            float qs2 = qsat(T, pfl);
            if (qv < qs2) {
                float rc = 0.001f * r;
                float rq = rc * qr;
                float C = 1.6f + 124.9f * asphm::fpow(rq, 0.2046f);
                float E = (((1.0f - qv / qs2) * C) * asphm::fpow(rq, 0.525f)) /
                          ((5.4e5f + 2.55e8f / (pfl * qs2)) * rc);
                float dqe = fminf(fminf(E * A.dt, qr), qs2 - qv);
                qv += dqe;
                th -= (L_V * dqe) / (CP * Pi);
                dfx -= to_fx(dqe);
            }
        }
    }
    if (dfx != 0) {
        if (cid < 0) cid = K.p0 + cell_of(P, pi.x, pi.y, pi.z);
        atomicAdd((unsigned long long*)&A.dep[cid], (unsigned long long)dfx);
    }
    th = fminf(fmaxf(th, 100.0f), P.theta_cap);
    qv = fminf(fmaxf(qv, 0.0f), P.qv_cap);
    qc = fmaxf(qc, 0.0f);
    A.thq[i] = make_float4(th, qv, qc, 0.f);
    A.prs[i] = ppr;
    reinterpret_cast<float*>(A.vel)[4 * (size_t)i + 3] = ppr;
}

// momentum finish (:497-510) + theta update, second kick, sponge (mode 1). Returns |v| maxima.
__device__ __forceinline__ void mom_epilogue(const DevParams& P, const MomArgs& A, int i, float4 pi, float4 vi,
                             float4 ti, float pri, MomAcc M, float& mx, float& my, float& mz) {
    float ax = M.ax * P.gmx, ay = M.ay * P.gmy, az = M.az * P.gmz;
    float avx = M.avx * P.gmx, avy = M.avy * P.gmy, avz = M.avz * P.gmz;
    const float z = pi.z;
    const float ex = exner_ref(P, z);
    const float thr = theta_ref(P, z);
    float thv = ti.x * (1.0f + 0.608f * ti.y);
    float thvr = thr * (1.0f + 0.608f * qv_ref(P, z));
    az = az + GRAV * ((thv - thvr) / thvr - ti.z);
    A.acc[i] = make_float4(ax, ay, az, 0.f);
    if (A.av_out) A.av_out[i] = make_float4(avx, avy, avz, 0.f);
    float Pi = asphm::fpow(fmaxf(pres_of_exner(P, ex) + pri, 1.0f) / P0, P.rd_cp);
    float dth = -__builtin_fmaf(vi.z, avz, __builtin_fmaf(vi.y, avy, vi.x * avx)) / (CP * Pi);
    dth = fminf(fmaxf(dth, -P.dthdt_cap), P.dthdt_cap);
    if (A.mode == 0) {
        A.dthdt[i] = dth;
        return;
    }
    float4 t = ti;
    t.x = fminf(fmaxf(__builtin_fmaf(dth, A.dt, t.x), 100.0f), P.theta_cap);
    A.thq[i] = t;
    const float hdt = 0.5f * A.dt;
    float4 v = vi;
    v.x = __builtin_fmaf(ax, hdt, v.x);
    v.y = __builtin_fmaf(ay, hdt, v.y);
    v.z = __builtin_fmaf(az, hdt, v.z);
    if ((P.flags & ASPH_F_SPONGE) && z >= P.sponge_zs) {
        float tt = (z - P.sponge_zs) / P.sponge_depth;
        float sg = (P.sponge_sigma * tt) * tt;
        v.z *= 1.0f / (1.0f + sg * A.dt);
    }
    v.w = 0.f;
    A.vel_out[i] = v;
    mx = fmaxf(mx, fabsf(v.x));
    my = fmaxf(my, fabsf(v.y));
    mz = fmaxf(mz, fabsf(v.z));
}

// ---------------------------------------------------------------------------------------------
// Brick staging shared by both kernels.
// ---------------------------------------------------------------------------------------------
template <int S, bool MOM> struct HaloMax {
    static constexpr int H = BS + 2 * (S + 1);
    static constexpr int HC = H * H * H;
    // halo particles held in LDS at a time: what the workgroup's share of the CU's LDS leaves
    // beside the cell table and the queues, at most 4032. Momentum (up to 128 VGPRs: 4 waves per
    // SIMD) takes 80 KiB, two workgroups per CU; density (fewer VGPRs) 52 KiB, three per CU.
    static constexpr int LDS = MOM ? 81920 : 53248;
    static constexpr int MAXP_FIT =
        ((LDS - (2 * (HC + 1) + 2 * queue_len<MOM>() * BT + 12 * H + 256)) / 16) / 64 * 64;
    static constexpr int MAXP = MAXP_FIT < 4032 ? MAXP_FIT : 4032;
};

template <int S, bool MOM> struct Smem {
    float4 p[HaloMax<S, MOM>::MAXP];   // x/dx, y/dy, z/dz (cells), and 1/max(rho_old,1e-3)
                                       // (density) or the particle's global index (momentum
                                       // and the reverse pair stages)
    uint16_t off[HaloMax<S, MOM>::HC + 1];   // exclusive start of each halo cell over the halo
    uint16_t q[MOM ? queue_len<MOM>() : 1][MOM ? BT : 1];   // per-thread queue of LDS slots (momentum)
    uint32_t mk[3][HaloMax<S, MOM>::H];  // Morton bits of the halo's x, y, z cells (wrapped);
                                         // z: ~0u where the layer lies outside [0, nz)
    uint32_t wsum[BW];
    uint16_t slab[NSLAB + 1];          // slab boundaries in halo rows (z * Hy + y)
    uint8_t slz[2][NSLAB];             // first and last halo z-layer of each slab
    int nslab;                         // 0: the halo cannot be staged (global-memory loop)
    int rowcut;                        // 1: slabs end inside z-layers (a layer exceeds MAXP)
    int modeA;                         // drift bound below 0.95 cell on every axis
    float delta[3];                    // drift bound since the last sort + rounding margin [cells]
};

struct BrickGeo {
    int bx0, by0, bz0;  // brick origin [cells]
    int hx0, hy0, hz0;  // halo origin
    int Hx, Hy, Hz;     // halo extent
    uint32_t mhx, mhl;  // ceil(2^32 / Hx), ceil(2^32 / (Hx Hy)): exact division of cell indices
    bool wraps;         // the halo crosses a periodic face
};

// Reads the drift bound since the last sort (all members) and whether it allows mode A.
template <int S, bool MOM>
__device__ __forceinline__ void prologue(const float* disp, const uint32_t* disp_slots, Smem<S, MOM>& sm) {
    const int tid = threadIdx.x;
    if (tid < 64) {
        float m[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float v = __uint_as_float(disp_slots[a * 64 + tid]);
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
            m[a] = v;
        }
        if (tid == 0) {
            bool ok = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float d = disp[a] + m[a];
                sm.delta[a] = d + DELTA_EPS;
                ok = ok && (d < 0.95f);
            }
            sm.modeA = ok ? 1 : 0;
        }
    }
    __syncthreads();
}

// The drift bound of one work item and whether it allows mode A: its brick's entry of the per-brick
// halo bounds (k_brick_halo), or without them the member-wide bound since the sort (prologue).
// A non-finite entry (a non-finite position, or a domain narrower than two bricks) gives mode B.
template <int S, bool MOM>
__device__ __forceinline__ void item_prologue(const float4* bhalo, int brick, const float* disp,
                                              const uint32_t* disp_slots, Smem<S, MOM>& sm) {
    if (!bhalo) {
        prologue<S, MOM>(disp, disp_slots, sm);
        return;
    }
    if (threadIdx.x == 0) {
        const float4 e = bhalo[brick];
        sm.delta[0] = e.x + DELTA_EPS;
        sm.delta[1] = e.y + DELTA_EPS;
        sm.delta[2] = e.z + DELTA_EPS;
        sm.modeA = (e.x < 0.95f && e.y < 0.95f && e.z < 0.95f) ? 1 : 0;
    }
    __syncthreads();
}

// Halo margin on one axis for a member: a particle lies within [c - d, c + 1 + d] of its listed
// cell c (d the drift bound), its current cell within c +- 1, and the cells its culled walk
// visits within c +- min(S + 1, floor(1 + 2d + R)), R the culling radius.
template <int S>
__device__ __forceinline__ int margin_of(const DevParams& P, float d) {
    const float R = sqrtf(P.lim * 1.0001f) + DELTA_EPS;
    const int mg = (int)floorf(1.0f + 2.0f * d + R);
    return mg < S + 1 ? mg : S + 1;
}

template <int S, bool MOM>
__device__ __forceinline__ BrickGeo brick_geo(const DevParams& P, const Smem<S, MOM>& sm, uint32_t kb) {
    BrickGeo g;
    g.bx0 = (int)compact1by2(kb);
    g.by0 = (int)compact1by2(kb >> 1);
    g.bz0 = (int)compact1by2(kb >> 2);
    const int mx = margin_of<S>(P, sm.delta[0]), my = margin_of<S>(P, sm.delta[1]),
              mz = margin_of<S>(P, sm.delta[2]);
    g.Hx = BS + 2 * mx;
    g.Hy = BS + 2 * my;
    g.Hz = BS + 2 * mz;
    g.hx0 = g.bx0 - mx;
    g.hy0 = g.by0 - my;
    g.hz0 = g.bz0 - mz;
    g.mhx = 0xffffffffu / (uint32_t)g.Hx + 1u;
    g.mhl = 0xffffffffu / (uint32_t)(g.Hx * g.Hy) + 1u;
    g.wraps = (g.hx0 < 0) || (g.hx0 + g.Hx > P.nx) || (g.hy0 < 0) || (g.hy0 + g.Hy > P.ny);
    return g;
}

// Key of halo cell c (x fastest, then y, then z). c and the divisors stay below 2^16, where
// floor(c * ceil(2^32 / d) / 2^32) = floor(c / d) exactly.
template <int S, bool MOM>
__device__ __forceinline__ uint32_t halo_key(uint32_t k0, const BrickGeo& g, const Smem<S, MOM>& sm, int c,
                                             bool& inside) {
    const uint32_t hz = __umulhi((uint32_t)c, g.mhl);
    const uint32_t rxy = (uint32_t)c - hz * (uint32_t)(g.Hx * g.Hy);
    const uint32_t hy = __umulhi(rxy, g.mhx);
    const uint32_t hx = rxy - hy * (uint32_t)g.Hx;
    const uint32_t kz = sm.mk[2][hz];
    inside = kz != ~0u;
    return k0 | sm.mk[0][hx] | sm.mk[1][hy] | (inside ? kz : 0u);
}

// Fills the Morton tables of the halo's cell coordinates (x and y wrapped periodically).
template <int S, bool MOM>
__device__ __forceinline__ void halo_tables(const DevParams& P, const BrickGeo& g, Smem<S, MOM>& sm) {
    static_assert(BT >= 192 && HaloMax<S, MOM>::H <= 64, "one thread per table entry");
    const int t = threadIdx.x & 63, a = threadIdx.x >> 6;
    if (a == 0 && t < g.Hx) sm.mk[0][t] = part1by2((uint32_t)wrapi(g.hx0 + t, P.nx));
    if (a == 1 && t < g.Hy) sm.mk[1][t] = part1by2((uint32_t)wrapi(g.hy0 + t, P.ny)) << 1;
    if (a == 2 && t < g.Hz) {
        const int gz = g.hz0 + t;
        sm.mk[2][t] = (gz >= 0 && gz < P.nz) ? part1by2((uint32_t)gz) << 2 : ~0u;
    }
    __syncthreads();
}

// Counts the halo's particles per cell, turns the counts into start offsets over the whole
// halo, and splits the halo into slabs of consecutive halo rows (z-major, then y) holding at most
// MAXP particles each. Sets sm.nslab = 0 if that is impossible (16-bit offsets, one row beyond
// MAXP, or more than NSLAB slabs).
template <int S, bool MOM>
__device__ __forceinline__ void plan(const DevParams& P, uint32_t k0, const BrickGeo& g,
                                     const uint32_t* __restrict__ cstart, Smem<S, MOM>& sm) {
    halo_tables<S, MOM>(P, g, sm);
    const int tid = threadIdx.x;
    const int HC = g.Hx * g.Hy * g.Hz;
    const int PER = (HC + BT - 1) / BT;
    const int c0 = tid * PER;
    uint32_t local = 0;
    for (int k = 0; k < PER; ++k) {
        const int c = c0 + k;
        if (c >= HC) break;
        bool inside;
        const uint32_t key = halo_key(k0, g, sm, c, inside);
        uint32_t n = inside ? cstart[key + 1] - cstart[key] : 0u;
        sm.off[c] = (uint16_t)(n < 65535u ? n : 65535u);
        local += n;
    }
    const int lane = tid & 63, w = tid >> 6;
    uint32_t x = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sm.wsum[w] = x;
    __syncthreads();
    uint32_t woff = 0, total = 0;
#pragma unroll
    for (int k = 0; k < BW; ++k) {
        uint32_t v = sm.wsum[k];
        if (k < w) woff += v;
        total += v;
    }
    const bool fits = total < 65535u;
    uint32_t run = woff + x - local;
    for (int k = 0; k < PER; ++k) {
        const int c = c0 + k;
        if (c >= HC) break;
        const uint32_t v = sm.off[c];
        sm.off[c] = (uint16_t)(fits ? run : 0u);
        run += v;
    }
    if (tid == 0) sm.off[HC] = (uint16_t)(fits ? total : 0u);
    const int R = g.Hz * g.Hy;   // halo rows
    if (fits && total <= (uint32_t)HaloMax<S, MOM>::MAXP) {   // workgroup-uniform: one slab
        if (tid == 0) {
            sm.slab[0] = 0;
            sm.slab[1] = (uint16_t)R;
            sm.slz[0][0] = 0;
            sm.slz[1][0] = (uint8_t)(g.Hz - 1);
            sm.rowcut = 0;
            sm.nslab = 1;
        }
        __syncthreads();
        return;
    }
    __syncthreads();
    if (tid == 0) {
        // greedy: each slab takes the most whole z-layers that fit (rowcut = 0), or where one
        // layer alone exceeds MAXP the most rows (rowcut = 1); a slab's end is found by bisection
        // over the layer or row starts (off[r * Hx] is nondecreasing in r)
        const uint32_t cap = (uint32_t)HaloMax<S, MOM>::MAXP;
        int ns = 0;
        bool ok = fits;
        int rowcut = 0;
        for (int pass = 0; pass < 2 && ok; ++pass) {
            const int unit = pass == 0 ? g.Hy : 1;   // rows per packing unit
            const int U = R / unit;
            ns = 0;
            int r = 0;
            sm.slab[0] = 0;
            bool done = true;
            while (r < U) {
                const uint32_t b = sm.off[r * unit * g.Hx];
                int lo = r, hi = U;   // the last unit end with at most MAXP particles after b
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if ((uint32_t)sm.off[mid * unit * g.Hx] - b <= cap) lo = mid;
                    else hi = mid - 1;
                }
                if (lo == r || ns == NSLAB) { done = false; break; }
                sm.slz[0][ns] = (uint8_t)(r * unit / g.Hy);
                sm.slz[1][ns] = (uint8_t)((lo * unit - 1) / g.Hy);
                sm.slab[++ns] = (uint16_t)(lo * unit);
                r = lo;
            }
            if (done) break;
            if (pass == 1 || ns == NSLAB) ok = false;   // one row beyond MAXP, or too many slabs
            rowcut = 1;
        }
        sm.rowcut = rowcut;
        sm.nslab = ok ? ns : 0;
    }
    __syncthreads();
}

// Stages slab s (its halo cells' particles, listed order) into LDS. Workgroup-uniform; waits
// for every wavefront to finish with the previous slab first.
template <int S, bool MOM, class Rec>
__device__ __forceinline__ void stage_slab(const DevParams& P, uint32_t k0, const BrickGeo& g, int s,
                                           const float4* __restrict__ pos, const uint32_t* __restrict__ cstart,
                                           Smem<S, MOM>& sm, Rec rec) {
    __syncthreads();
    const int cb = sm.slab[s] * g.Hx, ce = sm.slab[s + 1] * g.Hx;
    const uint32_t base = sm.off[cb];
    for (int c = cb + (int)threadIdx.x; c < ce; c += BT) {
        const uint32_t o = sm.off[c], n = (uint32_t)sm.off[c + 1] - o;
        if (n == 0) continue;
        bool inside;
        const uint32_t key = halo_key(k0, g, sm, c, inside);
        const uint32_t st = cstart[key];
        for (uint32_t m = 0; m < n; ++m) rec(sm, o - base + m, st + m, pos[st + m]);
    }
    __syncthreads();
}

// Current cell and cell-unit position of a particle in the brick's frame: a particle listed in
// the brick may have crossed a periodic face since the sort; its cell and position are shifted
// by a period so that its stencil rows index the halo (the pair arithmetic keeps the stored
// coordinates, as the reference does).
struct Cur {
    int ix, iy, z0, z1;
    float xi, yi, zi;
};

template <int S>
__device__ __forceinline__ Cur cur_cell(const DevParams& P, const BrickGeo& g, float4 pi) {
    Cur c;
    int iz;
    cell_trunc(P, pi.x, pi.y, pi.z, c.ix, c.iy, iz);
    zrange<S>(P, iz, c.z0, c.z1);
    c.xi = pi.x * P.inv_dx;
    c.yi = pi.y * P.inv_dy;
    c.zi = pi.z * P.inv_dz;
    // listed in [b0, b0 + BS) and at most one cell away from it: the offset lies in [-1, BS]
    const int dx = c.ix - g.bx0, dy = c.iy - g.by0;
    if (dx > BS) { c.ix -= P.nx; c.xi -= (float)P.nx; }
    else if (dx < -1) { c.ix += P.nx; c.xi += (float)P.nx; }
    if (dy > BS) { c.iy -= P.ny; c.yi -= (float)P.ny; }
    else if (dy < -1) { c.iy += P.ny; c.yi += (float)P.ny; }
    return c;
}

// Walks the culled stencil rows of a particle that lie in slab s: calls cand(slot) for every
// candidate in the reference order. Returns the number of candidates.
// ez, ey are the distances from the particle to a layer's or row's cells widened by the drift
// bound, [j - d, j + 1 + d]; rows beyond the support are skipped and the cells of a row are
// clipped to its chord. These are culling bounds only: their rounding (a few ulps of coordinates
// below 2^16 cells) and the 1-ulp hardware square root are covered by the 1e-3-cell margin, and
// every surviving candidate takes the exact support test. All lanes step through the same stencil
// offsets together (rows of equal offset hold similar spans), skipping rows individually; rows
// outside slab s are skipped the same way when the slab is part of the halo.
template <int S, bool ROWS, int G = 2, class Cand, class Sm>
__device__ __forceinline__ int walk_rows(const DevParams& P, const BrickGeo& g, const float* delta,
                                         const Sm& sm, int s, const Cur& c, Cand cand) {
    const float limc = P.lim * 1.0001f;
    const uint16_t* off = sm.off;
    // the slab's halo rows [r0, r1) hold its slots [0, cnt). With ROWS (slabs that end inside
    // z-layers) the rows of its first and last layers outside it clamp to an empty span.
    const int r0 = sm.slab[s], r1 = sm.slab[s + 1];
    const int base = __builtin_amdgcn_readfirstlane((int)off[r0 * g.Hx]);
    const int cnt = __builtin_amdgcn_readfirstlane((int)off[r1 * g.Hx]) - base;
    const int za = max(c.z0, g.hz0 + (int)sm.slz[0][s]);
    const int zb = min(c.z1, g.hz0 + (int)sm.slz[1][s]);
    const int ya = max(c.iy - S, g.hy0), yb = min(c.iy + S, g.hy0 + g.Hy - 1);
    const int xa = max(c.ix - S, g.hx0), xb = min(c.ix + S, g.hx0 + g.Hx - 1);
    // a cell j's widened box lies beyond the particle by j - zP on one side, zM - j on the other
    const float zP = c.zi + delta[2], zM = (c.zi - delta[2]) - 1.0f;
    const float yP = c.yi + delta[1], yM = (c.yi - delta[1]) - 1.0f;
    const float xP = c.xi + delta[0], xM = (c.xi - delta[0]) - 1.0f;
    const float fya = (float)ya;
    int ncand = 0;
    for (int jz = za; jz <= zb; ++jz) {
        const float fz = (float)jz;
        const float ez = fmaxf(0.f, fmaxf(fz - zP, zM - fz));
        const float ez2 = ez * ez;
        if (ez2 >= limc) continue;
        int row = ((jz - g.hz0) * g.Hy + (ya - g.hy0)) * g.Hx - g.hx0;
        float fy = fya;
        for (int jy = ya; jy <= yb; ++jy, fy += 1.0f, row += g.Hx) {
            const float ey = fmaxf(0.f, fmaxf(fy - yP, yM - fy));
            const float d2 = __builtin_fmaf(ey, ey, ez2);
            if (d2 >= limc) continue;
            const float ext = __builtin_amdgcn_sqrtf(limc - d2) + DELTA_EPS;
            const int xlo = max(xa, (int)ceilf(xM - ext)), xhi = min(xb, (int)floorf(xP + ext));
            if (xlo > xhi) continue;
            int a = (int)off[row + xlo] - base, e = (int)off[row + xhi + 1] - base;
            if (ROWS) {
                a = min(max(a, 0), cnt);
                e = min(max(e, 0), cnt);
            }
            ncand += e - a;
            // G candidates per step, fewer at the end of the span: cand(p, n), n = min(G, e - p);
            // G = 1: cand(q) with q the staged candidate's slot, one address stepping the span
            if constexpr (G == 1) {
                for (const float4 *q = sm.p + a, *qe = sm.p + e; q < qe; ++q) cand(q);
            } else {
                for (int p = a; p < e; p += G) cand(p, e - p);
            }
        }
    }
    return ncand;
}

// Adds 64 x the largest per-lane count of the wavefront (the lane-slots the wavefront spends) to
// a work counter.
__device__ __forceinline__ void count_wave_max(unsigned long long* stats, int idx, int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&stats[idx * 64 + (blockIdx.x & 63)], 64ull * (unsigned)v);
}

// Adds a per-lane count, summed over the wavefront, to a work counter.
__device__ __forceinline__ void count_wave(unsigned long long* stats, int idx, int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(&stats[idx * 64 + (blockIdx.x & 63)], (unsigned long long)v);
}

__device__ __forceinline__ void count_wg(unsigned long long* stats, int idx, int v = 1) {
    if (stats && threadIdx.x == 0) atomicAdd(&stats[idx * 64 + (blockIdx.x & 63)], (unsigned long long)v);
}

// ---------------------------------------------------------------------------------------------
// Density + EOS + microphysics.
// ---------------------------------------------------------------------------------------------
template <int S, bool WRAP, bool ROWS, class Rec>
__device__ __forceinline__ void dens_brick(const DevParams& P, const DensArgs& A, const DensConst& K,
                                           uint32_t k0, const BrickGeo& g, uint32_t o0, uint32_t o1,
                                           Smem<S, false>& sm, Rec rec) {
    const int tid = threadIdx.x;
    const int ns = sm.nslab;
    if (ns == 1) stage_slab<S, false>(P, k0, g, 0, A.pos, A.cstart, sm, rec);
    for (uint32_t r = o0; r < o1; r += BT) {
        const uint32_t i = r + tid;
        const bool active = i < o1;
        const float4 pi = active ? A.pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        const Cur cc = cur_cell<S>(P, g, pi);
        const float4 ci = cellpos(P, pi);
        float S1 = 0.f, S2 = 0.f;
        int npair = 0, ncand = 0;
        auto diff = [&](float4 cj, float& dxm, float& dym, float& dzm) { csep<WRAP>(P, ci, cj, dxm, dym, dzm); };
        for (int s = 0; s < ns; ++s) {
            if (ns > 1) stage_slab<S, false>(P, k0, g, s, A.pos, A.cstart, sm, rec);
            if (active) {
                // one candidate per step, in the reference order. The kernel value is
                // computed under the reference's test r2 < 4h^2 and added there, with no queue:
                // the square root and the spline are a branch that a wavefront skips when none
                // of its lanes has its candidate in support, as at the stencil's outer rows,
                // which all lanes walk together.
                ncand += walk_rows<S, ROWS, 1>(P, g, sm.delta, sm, s, cc, [&](const float4* q) {
                    const float4 pj = *q;
                    float dxm, dym, dzm;
                    diff(pj, dxm, dym, dzm);
                    const float r2 = r2of(dxm, dym, dzm);
                    const bool a = r2 < P.lim;
                    // outside the support the sums take +0, which leaves them unchanged
                    const float W = a ? fker_sel(sqrt_rn(r2) * P.inv_h) : 0.f;   // a branch
                    S1 += W;
                    S2 += W * pj.w;
                    npair += a ? 1 : 0;
                });
            }
        }
        if (active) dens_epilogue(P, A, K, (int)i, pi, S1, S2);
        if (A.stats) {
            count_wave(A.stats, ASPH_C_DENS_CAND, ncand);
            count_wave(A.stats, ASPH_C_DENS_PAIRS, npair);
            count_wave_max(A.stats, ASPH_C_DENS_WAVE_CAND, ncand);
        }
    }
}

// Whether a member's bricks take mode A (LDS) this launch: the drift bound and the halo width.
template <int S, bool MOM>
__device__ __forceinline__ bool member_mode_a(const DevParams& P, const Smem<S, MOM>& sm) {
    return sm.modeA && P.nx >= BS + 2 * margin_of<S>(P, sm.delta[0]) &&
           P.ny >= BS + 2 * margin_of<S>(P, sm.delta[1]);
}

// Owned range of work item `it` (brick it.x, chunk it.y): a brick's listed particles are split
// into equal chunks of at most `csize`.
__device__ __forceinline__ void item_range(const uint32_t* __restrict__ cstart, uint32_t k0, uint32_t kb, int chunk,
                                           uint32_t csize, uint32_t& o0, uint32_t& o1) {
    const uint32_t b0 = cstart[k0 + kb], b1 = cstart[k0 + kb + (1u << NB)];
    const uint32_t n = b1 - b0, ni = (n + csize - 1) / csize;
    const uint32_t q = ni ? (n + ni - 1) / ni : 0u;
    o0 = b0 + (uint32_t)chunk * q;
    o1 = min(b1, o0 + q);
}

// Work item taken by workgroup b of n. Workgroups go to the 8 XCDs in turn (b mod 8), each XCD
// with its own L2; within every group of 8 x XCD_RUN workgroups, XCD x takes the XCD_RUN
// consecutive items x * XCD_RUN + [0, XCD_RUN), bricks that are neighbours in key order and
// share most of their halos. The list order (heaviest first) holds from group to group. A
// bijection on [0, n): the last, partial group keeps the identity.
constexpr uint32_t XCD_RUN = 32;
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t n) {
    constexpr uint32_t G = 8 * XCD_RUN;
    if (b >= n / G * G) return b;
    const uint32_t j = b % G;
    return b - j + (j % 8) * XCD_RUN + j / 8;
}

// Workgroup b takes work item b: mode A (LDS halo) when the item's drift bound allows it and the
// domain is wider than its halo, else mode B over the item's particles. Every listed particle
// belongs to exactly one item.
template <int S>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(6))) void k_density_eos_brick(
        const MemberDesc* __restrict__ mem, const BlockMap* __restrict__ bmap, const int2* __restrict__ items,
        const uint32_t* __restrict__ nitems, int n_members, int csize, DensArgs A0) {
    __shared__ Smem<S, false> sm;
    static_assert(sizeof(Smem<S, false>) <= HaloMax<S, false>::LDS, "density LDS budget");
    DensArgs A = A0;
    if (A.ctl) {   // device controller: its dt; an idle step only rotates the density buffers
        if (A.ctl->idle) {
            const int64_t n = A.ctl->n;
            for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
                A.rho_new[i] = A.rho_old[i];
            return;
        }
        A.dt = A.ctl->dt;
    }
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    if (b < n_members && tid < 3 * 64) A.maxv[b * 192 + tid] = 0u;   // max|v| slots of member b
    const uint32_t nit = *nitems;
    if ((uint32_t)b >= nit) return;
    const int2 it = items[xcd_item((uint32_t)b, nit)];
    const BlockMap bm = bmap[it.x];
    const MemberDesc& MD = mem[bm.m];
    const DevParams& P = MD.P;
    const uint32_t k0 = MD.k0;
    const uint32_t kb = (uint32_t)bm.base;
    item_prologue<S, false>(A.bhalo, it.x, A.disp, A.disp_slots, sm);
    DensConst K;
    K.p0 = MD.p0;
    K.auto_f = auto_factor(P, A.dt);
    sfc_params(P, A.clock ? A.clock[0] : A.simt, K.Ts, K.qs_sfc);
    uint32_t o0, o1;
    item_range(A.cstart, k0, kb, it.y, (uint32_t)csize, o0, o1);
    if (o0 >= o1) return;
    const bool modeA = member_mode_a<S, false>(P, sm);
    if (modeA) {
        const BrickGeo g = brick_geo<S, false>(P, sm, kb);
        plan<S, false>(P, k0, g, A.cstart, sm);
    }
    if (!modeA || sm.nslab == 0) {
        // mode B (the item's drift bound, or a domain narrower than its halo), or a halo that
        // cannot be staged: thread per particle over the global cell table
        count_wg(A.stats, modeA ? ASPH_C_DENS_GLOBAL : ASPH_C_DENS_MODEB);
        for (uint32_t i = o0 + tid; i < o1; i += BT) {
            float4 pi = A.pos[i];
            float S1, S2;
            dens_sum_global<S>(P, k0, pi, A.pos, A.rho_old, A.cstart, S1, S2);
            dens_epilogue(P, A, K, (int)i, pi, S1, S2);
        }
        return;
    }
    const BrickGeo g = brick_geo<S, false>(P, sm, kb);
    count_wg(A.stats, sm.nslab == 1 ? ASPH_C_DENS_ONE_SLAB : ASPH_C_DENS_SLABBED);
    count_wg(A.stats, ASPH_C_DENS_STAGINGS, sm.nslab * (int)((o1 - o0 + BT - 1) / BT));
    const float* rho_old = A.rho_old;
    auto rec = [&](Smem<S, false>& s, uint32_t slot, uint32_t jg, float4 pj) {
        s.p[slot] = make_float4(pj.x * P.inv_dx, pj.y * P.inv_dy, pj.z * P.inv_dz, 1.0f / fmaxf(rho_old[jg], 1e-3f));
    };
    if (sm.rowcut) {
        if (g.wraps) dens_brick<S, true, true>(P, A, K, k0, g, o0, o1, sm, rec);
        else dens_brick<S, false, true>(P, A, K, k0, g, o0, o1, sm, rec);
    } else {
        if (g.wraps) dens_brick<S, true, false>(P, A, K, k0, g, o0, o1, sm, rec);
        else dens_brick<S, false, false>(P, A, K, k0, g, o0, o1, sm, rec);
    }
}

// ---------------------------------------------------------------------------------------------
// Momentum + theta update + kick + sponge + max|v|.
// ---------------------------------------------------------------------------------------------
template <int S, bool WRAP, bool ROWS, class Rec>
__device__ __forceinline__ void mom_brick(const DevParams& P, const MomArgs& A, uint32_t k0, const BrickGeo& g,
                                          uint32_t o0, uint32_t o1, Smem<S, true>& sm, Rec rec,
                                          float& mx, float& my, float& mz) {
    const int tid = threadIdx.x;
    const int ns = sm.nslab;
    if (ns == 1) stage_slab<S, true>(P, k0, g, 0, A.pos, A.cstart, sm, rec);
    for (uint32_t r = o0; r < o1; r += BT) {
        const uint32_t i = r + tid;
        const bool active = i < o1;
        float4 pi = make_float4(0.f, 0.f, 0.f, 0.f), vi = pi, ti = pi;
        float rhoi = 1.f, pri = 0.f;
        if (active) {
            pi = A.pos[i];
            vi = A.vel[i];
            ti = A.thq[i];
            rhoi = fmaxf(A.rho[i], 1e-3f);
            pri = A.prs[i];
        }
        const Cur cc = cur_cell<S>(P, g, pi);
        const float4 ci = cellpos(P, pi);
        const float inv_rhoi = 1.0f / rhoi;
        MomAcc M = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        int qn = 0, npair = 0, ncand = 0;
        auto diff = [&](float4 cj, float& dxt, float& dyt, float& dzt) { csep<WRAP>(P, ci, cj, dxt, dyt, dzt); };
        auto flush = [&]() {
            // lanes past their queue evaluate a placeholder (slot 0 of the staged slab, which
            // holds a particle whenever any lane has queued one) with zeroed factors, so there is
            // no divergent region. Called only when some lane has queued a pair; the loop tests
            // at the bottom, so the sums stay in place across iterations.
            MomAcc L = M;
            int k = 0;
            // MU queue entries per iteration: the operands of all its pairs are loaded before they
            // are evaluated (the sums still take them in queue order).
            do {
                bool act[MU];
                float4 pj[MU], vj[MU];   // vj = (vx, vy, vz, p'_j)
                float2 rj[MU];           // (max(rho_j, 1e-3), its reciprocal)
                uint32_t sl[MU];
                // slots read unconditionally (k + u < QK: k is a multiple of MU), so the reads are
                // in flight together; a stale entry is replaced by slot 0
#pragma unroll
                for (int u = 0; u < MU; ++u) sl[u] = sm.q[k + u][tid];
#pragma unroll
                for (int u = 0; u < MU; ++u) {
                    act[u] = k + u < qn;
                    pj[u] = sm.p[act[u] ? sl[u] : 0u];
                    vj[u] = A.vel[__float_as_uint(pj[u].w)];
                    rj[u] = A.rinv[__float_as_uint(pj[u].w)];
                }
#pragma unroll
                for (int u = 0; u < MU; ++u) {
                    if (u > 0 && !wave_any(act[u])) break;
                    float dxt, dyt, dzt;
                    diff(pj[u], dxt, dyt, dzt);
                    mom_pair(P, dxt, dyt, dzt, r2of(dxt, dyt, dzt), vi, rhoi, inv_rhoi, pri, rj[u].x, rj[u].y,
                             vj[u], vj[u].w, L, act[u]);
                }
                k += MU;
            } while (wave_any(k < qn));
            M = L;
            npair += qn;
            qn = 0;
        };
        for (int s = 0; s < ns; ++s) {
            if (ns > 1) stage_slab<S, true>(P, k0, g, s, A.pos, A.cstart, sm, rec);
            if (active) {
                // candidates p and p + 1 (if two): both staged positions are read before either
                // test; the queue takes them in order and is drained before it could overflow
                ncand += walk_rows<S, ROWS>(P, g, sm.delta, sm, s, cc, [&](int p, int n) {
                    const bool two = n > 1;
                    const float4 pj0 = sm.p[p], pj1 = sm.p[two ? p + 1 : p];
                    float dxt, dyt, dzt;
                    diff(pj0, dxt, dyt, dzt);
                    const float r20 = r2of(dxt, dyt, dzt);
                    diff(pj1, dxt, dyt, dzt);
                    const float r21 = r2of(dxt, dyt, dzt);
                    // both slots are written and the count takes the accepted ones: a rejected slot
                    // lies past the count and is overwritten (qn <= QK - 2 here)
                    sm.q[qn][tid] = (uint16_t)p;
                    qn += ((r20 < P.lim) & (r20 >= 1e-12f)) ? 1 : 0;
                    sm.q[qn][tid] = (uint16_t)(p + 1);
                    qn += (two & (r21 < P.lim) & (r21 >= 1e-12f)) ? 1 : 0;
                    if (wave_any(qn > QK - 2)) flush();
                });
            }
            if (wave_any(qn > 0)) flush();
        }
        if (active) mom_epilogue(P, A, (int)i, pi, vi, ti, pri, M, mx, my, mz);
        if (A.stats) {
            count_wave(A.stats, ASPH_C_MOM_CAND, ncand);
            count_wave(A.stats, ASPH_C_MOM_PAIRS, npair);
            count_wave_max(A.stats, ASPH_C_MOM_WAVE_CAND, ncand);
        }
    }
}

__device__ __forceinline__ void reduce_maxv(uint32_t* maxv, float mx, float my, float mz) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, o));
        my = fmaxf(my, __shfl_xor(my, o));
        mz = fmaxf(mz, __shfl_xor(mz, o));
    }
    if ((threadIdx.x & 63) == 0) {
        int slot = (blockIdx.x * BW + (threadIdx.x >> 6)) & 63;
        atomicMax(&maxv[slot], __float_as_uint(mx));
        atomicMax(&maxv[64 + slot], __float_as_uint(my));
        atomicMax(&maxv[128 + slot], __float_as_uint(mz));
    }
}

template <int S>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4))) void k_momentum_kdk_brick(
        const MemberDesc* __restrict__ mem, const BlockMap* __restrict__ bmap, const int2* __restrict__ items,
        const uint32_t* __restrict__ nitems, int csize, MomArgs A0) {
    __shared__ Smem<S, true> sm;
    static_assert(sizeof(Smem<S, true>) <= HaloMax<S, true>::LDS, "momentum LDS budget");
    MomArgs A = A0;
    if (A.ctl) {   // device controller: its dt; an idle step only rotates the velocity buffers
        if (A.ctl->idle) {
            const int64_t n = A.ctl->n;
            for (int64_t i = (int64_t)blockIdx.x * BT + threadIdx.x; i < n; i += (int64_t)gridDim.x * BT)
                A.vel_out[i] = A.vel[i];
            return;
        }
        A.dt = A.ctl->dt;
    }
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    const uint32_t nit = *nitems;
    if ((uint32_t)b >= nit) return;
    const int2 it = items[xcd_item((uint32_t)b, nit)];
    const BlockMap bm = bmap[it.x];
    const MemberDesc& MD = mem[bm.m];
    const DevParams& P = MD.P;
    const uint32_t k0 = MD.k0;
    const uint32_t kb = (uint32_t)bm.base;
    item_prologue<S, true>(A.bhalo, it.x, A.disp, A.disp_slots, sm);
    uint32_t o0, o1;
    item_range(A.cstart, k0, kb, it.y, (uint32_t)csize, o0, o1);
    if (o0 >= o1) return;
    float mx = 0.f, my = 0.f, mz = 0.f;
    const bool modeA = member_mode_a<S, true>(P, sm);
    if (modeA) {
        const BrickGeo g = brick_geo<S, true>(P, sm, kb);
        plan<S, true>(P, k0, g, A.cstart, sm);
    }
    if (!modeA || sm.nslab == 0) {
        // mode B (the item's drift bound, or a domain narrower than its halo), or a halo that
        // cannot be staged: thread per particle over the global cell table
        count_wg(A.stats, modeA ? ASPH_C_MOM_GLOBAL : ASPH_C_MOM_MODEB);
        for (uint32_t i = o0 + tid; i < o1; i += BT) {
            float4 pi = A.pos[i], vi = A.vel[i], ti = A.thq[i];
            float rhoi = fmaxf(A.rho[i], 1e-3f), pri = A.prs[i];
            MomAcc M = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            mom_sum_global<S>(P, k0, pi, vi, rhoi, pri, A.pos, A.vel, A.rho, A.prs, A.cstart, M);
            mom_epilogue(P, A, (int)i, pi, vi, ti, pri, M, mx, my, mz);
        }
    } else {
        const BrickGeo g = brick_geo<S, true>(P, sm, kb);
        count_wg(A.stats, sm.nslab == 1 ? ASPH_C_MOM_ONE_SLAB : ASPH_C_MOM_SLABBED);
        count_wg(A.stats, ASPH_C_MOM_STAGINGS, sm.nslab * (int)((o1 - o0 + BT - 1) / BT));
        auto rec = [&](Smem<S, true>& s, uint32_t slot, uint32_t jg, float4 pj) {
            s.p[slot] = make_float4(pj.x * P.inv_dx, pj.y * P.inv_dy, pj.z * P.inv_dz, __uint_as_float(jg));
        };
        if (sm.rowcut) {
            if (g.wraps) mom_brick<S, true, true>(P, A, k0, g, o0, o1, sm, rec, mx, my, mz);
            else mom_brick<S, false, true>(P, A, k0, g, o0, o1, sm, rec, mx, my, mz);
        } else {
            if (g.wraps) mom_brick<S, true, false>(P, A, k0, g, o0, o1, sm, rec, mx, my, mz);
            else mom_brick<S, false, false>(P, A, k0, g, o0, o1, sm, rec, mx, my, mz);
        }
    }
    if (A.mode) reduce_maxv(A.maxv + bm.m * 192, mx, my, mz);
}

// ---------------------------------------------------------------------------------------------
// Reverse (adjoint) pair stages on the staged halo.
//
// The reverse of particle i's momentum or density sum gathers, for i, the pair terms of (i, j) of
// i's own sum (own: j listed in the stencil of i's current cell) and of (j, i) of j's sum (rev: i
// listed in the stencil of j's current cell), each with the forward's support test. Every such
// j lies within the support of i, so (as in the forward walk) it is listed within
// floor(1 + 2d + R) cells of i's listed cell, d the drift bound since the sort and R the culling
// radius; by the stencils it is also listed within S + 1 + floor(d). While d < 1 the forward's
// halo (margin min(S + 1, floor(1 + 2d + R))) therefore holds every partner: each thread walks the
// halo cells within the margin of its listed cell, skipping cells whose box [c - d, c + 1 + d]
// lies outside the support (nearest periodic image), and queues in-support partners with their
// roles. Otherwise (d >= 1: mode B, by brick map), and for a halo that cannot be
// staged, each thread walks the global cell table within S + 1 + floor(d) cells.
// The pair terms (asph_adjoint.h) and the sets of pairs are those of the forward; the sums are
// float, in walk order.
// ---------------------------------------------------------------------------------------------
namespace adjrev {

// Whether cell (cx, cy, cz), 0 <= cz < nz, lies in the stencil of current cell (ix, iy, iz),
// asph_h100.cu:297-312. The z range clipped at the walls, [max(iz - S, 0), min(iz + S, nz - 1)],
// holds such a cz exactly when |cz - iz| <= S; across a periodic axis the offset d (|d| < n),
// folded into (-n/2, n/2], is at most S in magnitude exactly when |d| <= S or |d| >= n - S.
__device__ __forceinline__ bool near1(int d, int n, int S) {
    const int a = abs(d);
    return a <= S || a >= n - S;
}
__device__ __forceinline__ bool in_stencil(const DevParams& P, int ix, int iy, int iz, int cx, int cy, int cz) {
    const int S = P.S;
    return abs(cz - iz) <= S && near1(cx - ix, P.nx, S) && near1(cy - iy, P.ny, S);
}

// Distance [cells] from a point at offset t of a cell's centre to the cell's box widened by d,
// over the nearest periodic image when n > 0.
__device__ __forceinline__ float box_gap(float t, float d, int n) {
    if (n > 0) {
        const float hn = 0.5f * (float)n;
        if (t > hn) t -= (float)n;
        else if (t < -hn) t += (float)n;
    }
    return fmaxf(0.f, fabsf(t) - 0.5f - d);
}

struct Part {
    float xc, yc, zc;            // position [cells]
    int lx, ly, lz, ix, iy, iz;  // listed and current cell
};

__device__ __forceinline__ Part part(const DevParams& P, float4 x, uint32_t lkey) {
    Part a;
    a.xc = x.x * P.inv_dx;
    a.yc = x.y * P.inv_dy;
    a.zc = x.z * P.inv_dz;
    a.lx = (int)compact1by2(lkey);
    a.ly = (int)compact1by2(lkey >> 1);
    a.lz = (int)compact1by2(lkey >> 2);
    cell_trunc(P, x.x, x.y, x.z, a.ix, a.iy, a.iz);
    return a;
}

// The support test shared by both orientations (negating s is exact, so both square alike), from
// the coordinates in cells of i (ci) and j (cj).
__device__ __forceinline__ bool support(const DevParams& P, float4 ci, float4 cj) {
    float sx, sy, sz;
    csep(P, ci, cj, sx, sy, sz);
    return r2of(sx, sy, sz) < P.lim;
}

// Whether i is listed in the stencil of j's current cell. The cell is cell_trunc's, taken from j's
// coordinates in cells (the same products x * inv_dx).
__device__ __forceinline__ bool rev_role(const DevParams& P, float4 cj, const Part& a) {
    const int jx = min(max((int)cj.x, 0), P.nx - 1);
    const int jy = min(max((int)cj.y, 0), P.ny - 1);
    const int jz = min(max((int)cj.z, 0), P.nz - 1);
    return in_stencil(P, jx, jy, jz, a.lx, a.ly, a.lz);
}

// Global walk: f(j, own) for every particle listed within S + 1 + floor(d) cells of i's listed
// cell in a cell whose box meets the support.
template <class F>
__device__ __forceinline__ void walk_global(const DevParams& P, uint32_t k0, const uint32_t* __restrict__ cstart,
                                            const float* dl, const Part& a, F f) {
    const float limc = P.lim * 1.0001f;
    const int S = P.S;
    const int Mx = S + 1 + (int)dl[0], My = S + 1 + (int)dl[1], Mz = S + 1 + (int)dl[2];
    const int nxs = min(2 * Mx + 1, P.nx), nys = min(2 * My + 1, P.ny);
    for (int cz = max(a.lz - Mz, 0); cz <= min(a.lz + Mz, P.nz - 1); ++cz) {
        const float ez = box_gap(a.zc - ((float)cz + 0.5f), dl[2], 0);
        const float ez2 = ez * ez;
        if (ez2 >= limc) continue;
        for (int oy = 0; oy < nys; ++oy) {
            const int cyu = a.ly - My + oy;
            const float ey = box_gap(a.yc - ((float)cyu + 0.5f), dl[1], P.ny);
            const float eyz = __builtin_fmaf(ey, ey, ez2);
            if (eyz >= limc) continue;
            const int cy = wrapi(cyu, P.ny);
            for (int ox = 0; ox < nxs; ++ox) {
                const int cxu = a.lx - Mx + ox;
                const float ex = box_gap(a.xc - ((float)cxu + 0.5f), dl[0], P.nx);
                if (__builtin_fmaf(ex, ex, eyz) >= limc) continue;
                const int cx = wrapi(cxu, P.nx);
                const bool own = in_stencil(P, a.ix, a.iy, a.iz, cx, cy, cz);
                const uint32_t key = k0 | morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
                const uint32_t j0 = cstart[key], j1 = cstart[key + 1];
                for (uint32_t j = j0; j < j1; ++j) f(j, own);
            }
        }
    }
}

// Halo walk of slab s: f(slot, own) for the staged particles of the halo cells within the margin
// of i's listed cell (which lies in the brick) whose box [c - d, c + 1 + d] meets the support, in
// ascending (z, y, x). Per row the cells are culled to the chord of the support (as the forward
// walk does), in the brick's unwrapped frame: i's position and current cell are shifted by a period
// when i has crossed a periodic face since the sort. The window (2 S + 3 cells per axis, at most
// nx, ny) holds each cell once, so the unwrapped distances are the nearest-image ones.
template <int S, class F>
__device__ __forceinline__ void walk_halo(const DevParams& P, const BrickGeo& g, const Smem<S, true>& sm, int s,
                                          const Part& a, F f) {
    const float limc = P.lim * 1.0001f;
    const int mx = (g.Hx - BS) / 2, my = (g.Hy - BS) / 2, mz = (g.Hz - BS) / 2;
    const int hx = a.lx - g.hx0, hy = a.ly - g.hy0, hz = a.lz - g.hz0;
    const int r0 = sm.slab[s], r1 = sm.slab[s + 1];   // the slab's halo rows [r0, r1)
    const int base = sm.off[r0 * g.Hx];
    const int za = max(hz - mz, r0 / g.Hy), zb = min(hz + mz, (r1 - 1) / g.Hy);
    const int sx = a.ix - a.lx > P.nx / 2 ? -P.nx : (a.lx - a.ix > P.nx / 2 ? P.nx : 0);
    const int sy = a.iy - a.ly > P.ny / 2 ? -P.ny : (a.ly - a.iy > P.ny / 2 ? P.ny : 0);
    const float xu = a.xc + (float)sx, yu = a.yc + (float)sy;
    const int oxl = a.ix + sx - S - g.hx0, oxh = a.ix + sx + S - g.hx0;   // own x range (halo frame)
    const int iyu = a.iy + sy;
    for (int z = za; z <= zb; ++z) {
        const int cz = g.hz0 + z;
        if (cz < 0 || cz >= P.nz) continue;
        const float ez = box_gap(a.zc - ((float)cz + 0.5f), sm.delta[2], 0);
        const float ez2 = ez * ez;
        if (ez2 >= limc) continue;
        const bool oz = abs(cz - a.iz) <= S;
        for (int y = hy - my; y <= hy + my; ++y) {
            const int hr = z * g.Hy + y;
            if (hr < r0 || hr >= r1) continue;
            const int cyu = g.hy0 + y;
            const float ey = box_gap(yu - ((float)cyu + 0.5f), sm.delta[1], 0);
            const float eyz = __builtin_fmaf(ey, ey, ez2);
            if (eyz >= limc) continue;
            // a culling bound only: the 1e-3-cell margin covers the 1-ulp hardware square root
            const float ext = __builtin_amdgcn_sqrtf(limc - eyz) + DELTA_EPS;
            const int xl = max(hx - mx, (int)ceilf(xu - ext - 1.0f - sm.delta[0]) - g.hx0);
            const int xh = min(hx + mx, (int)floorf(xu + ext + sm.delta[0]) - g.hx0);
            if (xl > xh) continue;
            const bool oyz = oz && abs(cyu - iyu) <= S;
            const int row = (z * g.Hy + y) * g.Hx;
            // the row's cells xl..xh are one span of slots; its own cells (x in [oxl, oxh], when
            // the row is in the stencil of i's current cell) one contiguous part of it
            const int p0 = (int)sm.off[row + xl] - base, p1 = (int)sm.off[row + xh + 1] - base;
            const int ol = max(xl, oxl), oh = min(xh, oxh);
            int q0 = p1, q1 = p1;
            if (oyz && ol <= oh) {
                q0 = (int)sm.off[row + ol] - base;
                q1 = (int)sm.off[row + oh + 1] - base;
            }
            for (int p = p0; p < p1; ++p) f(p, p >= q0 && p < q1);
        }
    }
}

// The recorded drift bound of the step (with its rounding margin) and mode A (d < 1 per axis).
template <int S>
__device__ __forceinline__ void prologue(const float* __restrict__ delta, Smem<S, true>& sm) {
    if (threadIdx.x == 0) {
        bool ok = true;
        for (int c = 0; c < 3; ++c) {
            sm.delta[c] = delta[c];
            ok = ok && delta[c] < 1.0f;
        }
        sm.modeA = ok ? 1 : 0;
    }
    __syncthreads();
}

constexpr uint16_t QOWN = 0x8000, QREV = 0x4000, QSLOT = 0x3fff;   // LDS queue entry (slot < MAXP)

// ---- momentum --------------------------------------------------------------------------------
struct MomI {
    float4 xi, ci;   // position, and in cells
    float vi[3], rho_i, rho_raw, p_i, Ai[3], AVi[3];
    float xb[3], vb[3], rb, pb;
};

__device__ __forceinline__ MomI mom_load(const DevParams& P, const AdjTape& T, const AdjWork& W, uint32_t i) {
    MomI m;
    m.xi = T.x1[i];
    m.ci = cellpos(P, m.xi);
    const float4 v = T.v1[i], A = W.Ab[i], AV = W.AVb[i];
    m.vi[0] = v.x; m.vi[1] = v.y; m.vi[2] = v.z;
    m.rho_raw = T.r1[i];
    m.rho_i = fmaxf(m.rho_raw, 1e-3f);
    m.p_i = v.w;
    m.Ai[0] = A.x; m.Ai[1] = A.y; m.Ai[2] = A.z;
    m.AVi[0] = AV.x; m.AVi[1] = AV.y; m.AVi[2] = AV.z;
    for (int c = 0; c < 3; ++c) m.xb[c] = m.vb[c] = 0.f;
    m.rb = m.pb = 0.f;
    return m;
}

// The partner's operands: (v, p') and density of the step, and (rev) its seeds.
struct MomJ {
    float4 v, A, AV;
    float rho;
};
__device__ __forceinline__ MomJ mom_operands(const AdjTape& T, const AdjWork& W, uint32_t j, bool rev) {
    MomJ o;
    o.v = T.v1[j];
    o.rho = T.r1[j];
    // a pair outside one role enters with zero seeds for that orientation
    o.A = make_float4(0.f, 0.f, 0.f, 0.f);
    o.AV = o.A;
    if (rev) { o.A = W.Ab[j]; o.AV = W.AVb[j]; }
    return o;
}

__device__ __forceinline__ void mom_partner(const DevParams& P, MomI& m, float4 cj, const MomJ& o, bool own) {
    float sx, sy, sz;
    csep(P, m.ci, cj, sx, sy, sz);
    const float4 vj4 = o.v;
    const float vj[3] = {vj4.x, vj4.y, vj4.z};
    const float rho_j = fmaxf(o.rho, 1e-3f);
    const float Aj[3] = {o.A.x, o.A.y, o.A.z}, AVj[3] = {o.AV.x, o.AV.y, o.AV.z};
    const float z3[3] = {0.f, 0.f, 0.f};
    adj::mom_pair_vjp2<float>(P, sx, sy, sz, m.vi, vj, m.rho_i, rho_j, m.p_i, vj4.w, own ? m.Ai : z3,
                              own ? m.AVi : z3, Aj, AVj, m.xb, m.vb, m.rb, m.pb);
}

__device__ __forceinline__ void mom_store(const AdjWork& W, uint32_t i, const MomI& m) {
    const float4 x1b = W.x1b[i], v1b = W.v1b[i];
    W.x1b[i] = make_float4(x1b.x + m.xb[0], x1b.y + m.xb[1], x1b.z + m.xb[2], 0.f);
    W.v1b[i] = make_float4(v1b.x + m.vb[0], v1b.y + m.vb[1], v1b.z + m.vb[2], 0.f);
    if (m.rho_raw > 1e-3f) W.r1b[i] += m.rb;
    W.pb[i] += m.pb;
}

__device__ __forceinline__ void mom_global(const DevParams& P, uint32_t k0, const AdjTape& T, const AdjWork& W,
                                           const float* dl, uint32_t i) {
    MomI m = mom_load(P, T, W, i);
    const Part a = part(P, m.xi, T.keys[i] - k0);
    walk_global(P, k0, T.cstart, dl, a, [&](uint32_t j, bool own) {
        const float4 cj = cellpos(P, T.x1[j]);
        if (!support(P, m.ci, cj)) return;
        const bool rev = rev_role(P, cj, a);
        if (own || rev) mom_partner(P, m, cj, mom_operands(T, W, j, rev), own);
    });
    mom_store(W, i, m);
}

// ---- density ---------------------------------------------------------------------------------
struct DensI {
    float4 xi, ci;   // position, and in cells
    float r0_raw, inv_r0i;
    float2 sbi;
    float xb[3], rb;
};

__device__ __forceinline__ DensI dens_load(const DevParams& P, const AdjTape& T, const AdjWork& W, uint32_t i) {
    DensI d;
    d.xi = T.x1[i];
    d.ci = cellpos(P, d.xi);
    d.r0_raw = T.r0[i];
    d.inv_r0i = adj::Fn<float>::rcp(fmaxf(d.r0_raw, 1e-3f));
    d.sbi = W.sb[i];
    d.xb[0] = d.xb[1] = d.xb[2] = 0.f;
    d.rb = 0.f;
    return d;
}

// W and its gradient are even and odd in r, so one evaluation serves both orientations.
// The partner's operands: its start-of-step density (own) and density-sum seeds (rev).
struct DensJ {
    float r0;
    float2 sb;
};
__device__ __forceinline__ DensJ dens_operands(const AdjTape& T, const AdjWork& W, uint32_t j, bool own, bool rev) {
    DensJ o;
    o.r0 = own ? T.r0[j] : 1.f;
    o.sb = rev ? W.sb[j] : make_float2(0.f, 0.f);
    return o;
}

__device__ __forceinline__ void dens_partner(const DevParams& P, DensI& d, float4 cj, const DensJ& o, bool own,
                                             bool rev) {
    float sx, sy, sz;
    csep(P, d.ci, cj, sx, sy, sz);
    float Wv, gw[3];
    if (!adj::dens_pair<float>(P, sx, sy, sz, Wv, gw)) return;
    float wb = 0.f;
    if (own) wb = d.sbi.x + d.sbi.y * adj::Fn<float>::rcp(fmaxf(o.r0, 1e-3f));
    if (rev) {   // (j, i): r_ji = -r_ij, the gradient flips sign and x_i enters with a minus
        const float2 sbj = o.sb;
        wb += sbj.x + sbj.y * d.inv_r0i;
        d.rb += sbj.y * Wv * (-d.inv_r0i * d.inv_r0i);
    }
    for (int c = 0; c < 3; ++c) d.xb[c] += wb * gw[c];
}

__device__ __forceinline__ void dens_store(const AdjWork& W, const AdjState& Sst, uint32_t i, const DensI& d) {
    const float4 x1b = W.x1b[i];
    W.x1b[i] = make_float4(x1b.x + d.xb[0], x1b.y + d.xb[1], x1b.z + d.xb[2], 0.f);
    Sst.rho[i] = d.r0_raw > 1e-3f ? d.rb : 0.f;
}

__device__ __forceinline__ void dens_global(const DevParams& P, uint32_t k0, const AdjTape& T, const AdjWork& W,
                                            const AdjState& Sst, const float* dl, uint32_t i) {
    DensI d = dens_load(P, T, W, i);
    const Part a = part(P, d.xi, T.keys[i] - k0);
    walk_global(P, k0, T.cstart, dl, a, [&](uint32_t j, bool own) {
        const float4 cj = cellpos(P, T.x1[j]);
        if (!support(P, d.ci, cj)) return;
        const bool rev = rev_role(P, cj, a);
        if (own || rev) dens_partner(P, d, cj, dens_operands(T, W, j, own, rev), own, rev);
    });
    dens_store(W, Sst, i, d);
}

}  // namespace adjrev

// One kernel per pair stage, shaped as the forward's (work items of the segment's sort, mode B by
// brick map). KIND 0: momentum, 1: density. The adjoint runs on one member.
template <int S, int KIND>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(4))) void k_adj_pairs_brick(
        const MemberDesc* __restrict__ mem, const BlockMap* __restrict__ bmap, const int2* __restrict__ items,
        const uint32_t* __restrict__ nitems, int nbblk, int csize, AdjTape T, AdjWork W, AdjState Sst) {
    using namespace adjrev;
    __shared__ Smem<S, true> sm;
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    adjrev::prologue<S>(T.delta, sm);
    const MemberDesc& MD = mem[0];
    const DevParams& P = MD.P;
    const uint32_t k0 = MD.k0;
    auto one_global = [&](uint32_t i) {
        if (KIND == 0) mom_global(P, k0, T, W, sm.delta, i);
        else dens_global(P, k0, T, W, Sst, sm.delta, i);
    };
    if (!member_mode_a<S, true>(P, sm)) {
        if (b < nbblk) {
            const int il = bmap[b].ord * BT + tid;
            if (il < P.N) one_global((uint32_t)(MD.p0 + il));
        }
        return;
    }
    const uint32_t nit = *nitems;
    if ((uint32_t)b >= nit) return;
    const int2 it = items[xcd_item((uint32_t)b, nit)];
    const uint32_t kb = (uint32_t)bmap[it.x].base;
    uint32_t o0, o1;
    item_range(T.cstart, k0, kb, it.y, (uint32_t)csize, o0, o1);
    if (o0 >= o1) return;
    const BrickGeo g = brick_geo<S, true>(P, sm, kb);
    plan<S, true>(P, k0, g, T.cstart, sm);
    if (sm.nslab == 0) {
        for (uint32_t i = o0 + tid; i < o1; i += BT) one_global(i);
        return;
    }
    auto rec = [&](Smem<S, true>& s, uint32_t slot, uint32_t jg, float4 pj) {
        s.p[slot] = make_float4(pj.x * P.inv_dx, pj.y * P.inv_dy, pj.z * P.inv_dz, __uint_as_float(jg));
    };
    const int ns = sm.nslab;
    if (ns == 1) stage_slab<S, true>(P, k0, g, 0, T.x1, T.cstart, sm, rec);
    for (uint32_t r = o0; r < o1; r += BT) {
        const uint32_t i = r + tid;
        const bool active = i < o1;
        const uint32_t il = active ? i : o0;
        MomI m;
        DensI d;
        if (KIND == 0) m = mom_load(P, T, W, il);
        else d = dens_load(P, T, W, il);
        const float4 xi = KIND == 0 ? m.xi : d.xi;
        const float4 ci = KIND == 0 ? m.ci : d.ci;
        const Part a = part(P, xi, T.keys[il] - k0);
        int qn = 0;
        // Two queue entries per iteration: their staged positions (and, for density, the
        // partners' operands) are loaded before either is evaluated; the sums still take them in
        // queue order.
        auto flush = [&]() {
            for (int k = 0; wave_any(k < qn); k += 2) {
                const bool a0 = k < qn, a1 = k + 1 < qn;   // k + 1 < QK: k is even
                const uint16_t e0 = a0 ? sm.q[k][tid] : (uint16_t)0, e1 = a1 ? sm.q[k + 1][tid] : (uint16_t)0;
                const float4 pj0 = sm.p[e0 & QSLOT], pj1 = sm.p[e1 & QSLOT];
                const uint32_t j0 = __float_as_uint(pj0.w), j1 = __float_as_uint(pj1.w);
                const bool own0 = (e0 & QOWN) != 0, rev0 = (e0 & QREV) != 0;
                const bool own1 = (e1 & QOWN) != 0, rev1 = (e1 & QREV) != 0;
                if (KIND == 0) {   // (the momentum partner's registers leave room for one at a time)
                    if (a0) mom_partner(P, m, pj0, mom_operands(T, W, j0, rev0), own0);
                    if (a1) mom_partner(P, m, pj1, mom_operands(T, W, j1, rev1), own1);
                } else {
                    const DensJ o0 = dens_operands(T, W, j0, own0, rev0), o1 = dens_operands(T, W, j1, own1, rev1);
                    if (a0) dens_partner(P, d, pj0, o0, own0, rev0);
                    if (a1) dens_partner(P, d, pj1, o1, own1, rev1);
                }
            }
            qn = 0;
        };
        for (int s = 0; s < ns; ++s) {
            if (ns > 1) stage_slab<S, true>(P, k0, g, s, T.x1, T.cstart, sm, rec);
            if (active) {
                walk_halo<S>(P, g, sm, s, a, [&](int slot, bool own) {
                    const float4 cj = sm.p[slot];
                    if (support(P, ci, cj)) {
                        const bool rev = rev_role(P, cj, a);
                        if (own || rev)
                            sm.q[qn++][tid] = (uint16_t)((uint32_t)slot | (own ? QOWN : 0u) | (rev ? QREV : 0u));
                    }
                    if (wave_any(qn == QK)) flush();
                });
            }
            flush();
        }
        if (active) {
            if (KIND == 0) mom_store(W, i, m);
            else dens_store(W, Sst, i, d);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Work items of the brick kernels: every non-empty brick is split into chunks of at most `csize`
// listed particles, and (order = 1) the chunks are ordered by the brick's estimated cost (chunk
// size x brick occupancy, 4 bins per octave), heaviest first, so that the long workgroups of dense
// regions start early instead of forming the tail of the launch. Stable counting sort (count,
// scan, fill): within a bin, and for order = 0 over the whole map, the items keep the brick order
// (Morton order of each member), so that consecutive items, which run on one XCD, share halos.
// ---------------------------------------------------------------------------------------------
constexpr int ITEM_BINS = 128;
constexpr int ITEM_SEGS = 8;    // segments of the count workgroups per bin in k_items_scan

// Per-brick item count and cost bin, and the items of each bin in this workgroup's 256 bricks
// (hist[blockIdx.x][bin], every bin written).
__global__ __launch_bounds__(256) void k_items_count(const MemberDesc* __restrict__ mem,
                                                     const BlockMap* __restrict__ bmap, int nbblk,
                                                     uint32_t csize, int order,
                                                     const uint32_t* __restrict__ cstart,
                                                     uint32_t* __restrict__ meta, uint32_t* __restrict__ hist,
                                                     const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    __shared__ uint32_t lh[ITEM_BINS];
    const int t = threadIdx.x;
    if (t < ITEM_BINS) lh[t] = 0u;
    __syncthreads();
    const int b = blockIdx.x * 256 + t;
    if (b < nbblk) {
        const BlockMap bm = bmap[b];
        const uint32_t k0 = mem[bm.m].k0, kb = (uint32_t)bm.base;
        const uint32_t n = cstart[k0 + kb + (1u << NB)] - cstart[k0 + kb];
        const uint32_t ni = (n + csize - 1) / csize;
        int bin = 0;
        if (ni) {
            if (order) {
                const float cost = (float)min(n, csize) * (float)n;
                const int lg = (int)(4.0f * log2f(fmaxf(cost, 1.0f)));
                bin = ITEM_BINS - 1 - min(lg, ITEM_BINS - 1);
            }
            atomicAdd(&lh[bin], ni);
        }
        meta[b] = (ni << 8) | (uint32_t)bin;
    }
    __syncthreads();
    if (t < ITEM_BINS) hist[(size_t)blockIdx.x * ITEM_BINS + t] = lh[t];
}

// One workgroup: thread (segment j, bin k) sums bin k over its segment of the g count workgroups;
// the bins are scanned (heaviest first) and the segments within each bin, and each count is then
// replaced by the position of the workgroup's first item of that bin. Sets the item count.
__global__ __launch_bounds__(ITEM_BINS * ITEM_SEGS) void k_items_scan(uint32_t* __restrict__ hist, int g,
                                                                      uint32_t* __restrict__ nitems,
                                                                      const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    __shared__ uint32_t seg[ITEM_SEGS][ITEM_BINS];
    __shared__ uint32_t v[ITEM_BINS];
    const int k = threadIdx.x % ITEM_BINS, j = threadIdx.x / ITEM_BINS;
    const int per = (g + ITEM_SEGS - 1) / ITEM_SEGS;
    const int w0 = min(g, j * per), w1 = min(g, w0 + per);
    uint32_t tot = 0u;
#pragma unroll 8
    for (int w = w0; w < w1; ++w) tot += hist[(size_t)w * ITEM_BINS + k];
    seg[j][k] = tot;
    __syncthreads();
    if (j == 0) {
        uint32_t bin_tot = 0u;
        for (int q = 0; q < ITEM_SEGS; ++q) bin_tot += seg[q][k];
        v[k] = bin_tot;
    }
    __syncthreads();
    for (int o = 1; o < ITEM_BINS; o <<= 1) {   // inclusive scan of the bin totals
        uint32_t add = (j == 0 && k >= o) ? v[k - o] : 0u;
        __syncthreads();
        if (j == 0) v[k] += add;
        __syncthreads();
    }
    uint32_t run = k ? v[k - 1] : 0u;
    for (int q = 0; q < j; ++q) run += seg[q][k];
#pragma unroll 8
    for (int w = w0; w < w1; ++w) {
        const size_t a = (size_t)w * ITEM_BINS + k;
        const uint32_t c = hist[a];
        hist[a] = run;
        run += c;
    }
    if (threadIdx.x == ITEM_BINS - 1) *nitems = v[ITEM_BINS - 1];
}

// Writes each brick's items at its workgroup's position in its bin plus the items of the same bin
// from the lower bricks of the workgroup.
__global__ __launch_bounds__(256) void k_items_fill(int nbblk, const uint32_t* __restrict__ meta,
                                                    const uint32_t* __restrict__ hist, int2* __restrict__ items,
                                                    const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    __shared__ uint32_t lm[256];
    const int t = threadIdx.x;
    const int b = blockIdx.x * 256 + t;
    const uint32_t m = b < nbblk ? meta[b] : 0u;
    lm[t] = m;
    __syncthreads();
    const uint32_t ni = m >> 8;
    if (!ni) return;
    const uint32_t bin = m & 0xffu;
    uint32_t pos = hist[(size_t)blockIdx.x * ITEM_BINS + bin];
    for (int q = 0; q < t; ++q) {
        const uint32_t o = lm[q];
        pos += (o & 0xffu) == bin ? o >> 8 : 0u;
    }
    for (uint32_t k = 0; k < ni; ++k) items[pos + k] = make_int2(b, (int)k);
}

}  // namespace

int brick_bits() { return NB; }
size_t brick_item_hist_words(int nbblk) { return (size_t)((nbblk + 255) / 256) * ITEM_BINS; }
int brick_item_capacity(int nbblk, int64_t n) { return nbblk + (int)((n + CHUNK_MIN - 1) / CHUNK_MIN); }
int brick_chunk_min() { return CHUNK_MIN; }
int brick_chunk_max() { return CHUNK_MAX; }

void launch_brick_items(const Launch& L, hipStream_t s, const uint32_t* cstart) {
    const int g = (L.nbblk + 255) / 256;
    k_items_count<<<g, 256, 0, s>>>(L.mem, L.bmap, L.nbblk, (uint32_t)L.chunk, L.order, cstart, L.item_meta,
                                    L.item_hist, L.ctl);
    k_items_scan<<<1, ITEM_BINS * ITEM_SEGS, 0, s>>>(L.item_hist, g, L.nitems, L.ctl);
    k_items_fill<<<g, 256, 0, s>>>(L.nbblk, L.item_meta, L.item_hist, L.items, L.ctl);
}

// ---------------------------------------------------------------------------------------------
// Drift bound of each brick's work items: the largest per-brick bound (k_brick_exc) over the
// bricks that hold the cells of the brick's widest halo, brick +- (S + 1) cells, wrapped in x and
// y, clipped in z. Every particle an item of the brick owns, stages or walks is listed there. One
// thread per brick of the brick map. Where nx and ny are multiples of the brick edge those are the
// bricks at -BS, 0, +BS cells along each axis; otherwise they are enumerated cell by cell, and a
// domain narrower than two bricks, where more than 4 per axis could hold the halo, gets +inf
// (mode B).
// ---------------------------------------------------------------------------------------------
// The bricks (origins) holding the cells [c0 - S - 1, c0 + BS + S + 1) of one axis, in order along
// it; a brick met again after a wrap counts again. Returns how many there are; the first 4 land in
// b0..b3.
template <int S>
__device__ __forceinline__ int halo_bricks(int c0, int n, bool periodic, int& b0, int& b1, int& b2, int& b3) {
    int cnt = 0, prev = -1;
    b0 = b1 = b2 = b3 = c0;
#pragma unroll
    for (int d = -S - 1; d < BS + S + 1; ++d) {
        int c = c0 + d;
        if (periodic) {
            c = wrapi(c, n);
        } else if (c < 0 || c >= n) {
            continue;
        }
        const int bk = c & ~(BS - 1);
        if (bk != prev) {
            b0 = cnt == 0 ? bk : b0;
            b1 = cnt == 1 ? bk : b1;
            b2 = cnt == 2 ? bk : b2;
            b3 = cnt == 3 ? bk : b3;
            ++cnt;
            prev = bk;
        }
    }
    return cnt;
}

template <int S>
__global__ __launch_bounds__(256) void k_brick_halo(const MemberDesc* __restrict__ mem,
                                                    const BlockMap* __restrict__ bmap, int nbblk,
                                                    const float4* __restrict__ bexc, float4* __restrict__ bhalo,
                                                    const StepCtl* __restrict__ ctl) {
    if (ctl && ctl->idle) return;
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= nbblk) return;
    const BlockMap bm = bmap[g];
    const MemberDesc& M = mem[bm.m];
    const DevParams& P = M.P;
    const uint32_t kb = (uint32_t)bm.base;
    const int bx0 = (int)compact1by2(kb), by0 = (int)compact1by2(kb >> 1), bz0 = (int)compact1by2(kb >> 2);
    int x[4], y[4], z[4];
    int nx, ny, nz;
    if (((P.nx | P.ny) & (BS - 1)) == 0) {
        // the bricks at -BS, 0, +BS cells along each axis (S + 1 < BS): wrapped in x and y, the
        // ones inside [0, nz) in z
        x[0] = bx0 - BS < 0 ? bx0 - BS + P.nx : bx0 - BS;
        x[1] = bx0;
        x[2] = bx0 + BS >= P.nx ? bx0 + BS - P.nx : bx0 + BS;
        y[0] = by0 - BS < 0 ? by0 - BS + P.ny : by0 - BS;
        y[1] = by0;
        y[2] = by0 + BS >= P.ny ? by0 + BS - P.ny : by0 + BS;
        const bool zlo = bz0 >= BS, zhi = bz0 + BS < P.nz;
        z[0] = zlo ? bz0 - BS : bz0;
        z[1] = zlo ? bz0 : bz0 + BS;
        z[2] = bz0 + BS;
        x[3] = y[3] = z[3] = 0;
        nx = ny = 3;
        nz = 1 + (zlo ? 1 : 0) + (zhi ? 1 : 0);
    } else {
        nx = halo_bricks<S>(bx0, P.nx, true, x[0], x[1], x[2], x[3]);
        ny = halo_bricks<S>(by0, P.ny, true, y[0], y[1], y[2], y[3]);
        nz = halo_bricks<S>(bz0, P.nz, false, z[0], z[1], z[2], z[3]);
    }
    float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
    bool finite = true;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (a < nz && b < ny && c < nx) {
                    const float4 e = bexc[(M.k0 + morton3((uint32_t)x[c], (uint32_t)y[b], (uint32_t)z[a])) >> NB];
                    finite = finite && e.x == e.x && e.y == e.y && e.z == e.z;
                    m.x = fmaxf(m.x, e.x);
                    m.y = fmaxf(m.y, e.y);
                    m.z = fmaxf(m.z, e.z);
                }
    if (!finite || nx > 4 || ny > 4 || nz > 4) m.x = __int_as_float(0x7f800000);
    bhalo[g] = m;
}

void launch_brick_halo(const Launch& L, int S, hipStream_t s) {
    const int grid = (L.nbblk + 255) / 256;
    switch (S) {
        case 1: k_brick_halo<1><<<grid, 256, 0, s>>>(L.mem, L.bmap, L.nbblk, L.bexc, L.bhalo, L.ctl); break;
        case 2: k_brick_halo<2><<<grid, 256, 0, s>>>(L.mem, L.bmap, L.nbblk, L.bexc, L.bhalo, L.ctl); break;
        case 3: k_brick_halo<3><<<grid, 256, 0, s>>>(L.mem, L.bmap, L.nbblk, L.bexc, L.bhalo, L.ctl); break;
        default: k_brick_halo<4><<<grid, 256, 0, s>>>(L.mem, L.bmap, L.nbblk, L.bexc, L.bhalo, L.ctl); break;
    }
}

void launch_density_eos(const Launch& L, int S, hipStream_t s, const DensArgs& A) {
    dim3 g(L.nitem_max);
    switch (S) {
        case 1: k_density_eos_brick<1><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.n_members, L.chunk, A); break;
        case 2: k_density_eos_brick<2><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.n_members, L.chunk, A); break;
        case 3: k_density_eos_brick<3><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.n_members, L.chunk, A); break;
        default: k_density_eos_brick<4><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.n_members, L.chunk, A); break;
    }
}

void launch_momentum_kdk(const Launch& L, int S, hipStream_t s, const MomArgs& A) {
    dim3 g(L.nitem_max);
    switch (S) {
        case 1: k_momentum_kdk_brick<1><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.chunk, A); break;
        case 2: k_momentum_kdk_brick<2><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.chunk, A); break;
        case 3: k_momentum_kdk_brick<3><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.chunk, A); break;
        default: k_momentum_kdk_brick<4><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.chunk, A); break;
    }
}

void launch_adj_pairs(const Launch& L, int S, hipStream_t s, int kind, const AdjTape& T, const AdjWork& W,
                      const AdjState& Sst) {
    dim3 g(L.nitem_max);
#define ASPH_ADJ_PAIRS(SS)                                                                                        \
    if (kind == 0) k_adj_pairs_brick<SS, 0><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.nbblk, L.chunk, T, W, Sst); \
    else k_adj_pairs_brick<SS, 1><<<g, BT, 0, s>>>(L.mem, L.bmap, L.items, L.nitems, L.nbblk, L.chunk, T, W, Sst);
    switch (S) {
        case 1: ASPH_ADJ_PAIRS(1) break;
        case 2: ASPH_ADJ_PAIRS(2) break;
        case 3: ASPH_ADJ_PAIRS(3) break;
        default: ASPH_ADJ_PAIRS(4) break;
    }
#undef ASPH_ADJ_PAIRS
}

}  // namespace asph
