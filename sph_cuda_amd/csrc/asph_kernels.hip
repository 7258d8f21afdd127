// asph_kernels.hip - gfx950 kernels of one KDK step of the anisotropic-SPH model.
//
// Launch shape: one work-item per particle (or cell), 256 work-items (4 wave64) per
// workgroup. Particles are stored in Morton order after every sort, so the 64 particles of a
// wavefront occupy a 4x4x4 block of cells and their neighbour stencils overlap: the
// neighbour reads of a wavefront hit the same cache lines.
//
// Neighbour structure: cstart[k] = index of the first particle whose Morton key is >= k, for
// every key k in [0, 2^end_bit], so cell k holds particles [cstart[k], cstart[k+1]) -- one
// 4-byte start table in key order instead of the reference's int2 cellRange in row-major
// order (asph_h100.cu:257-278), written with coalesced stores and no clearing pass.
//
// Summation order inside every neighbour loop is the reference's jz -> jy -> jx -> j
// (asph_h100.cu:306-327, 444-495), so results do not depend on the launch shape.
#include "asph_kernels.h"

namespace asph {

namespace {

constexpr int TPB = 256;

__device__ __forceinline__ void min_image(const DevParams& P, float& rx, float& ry) {
    if (P.flags & ASPH_F_MIN_IMAGE) {
        if (rx > P.hLx) rx -= P.Lx; else if (rx < -P.hLx) rx += P.Lx;
        if (ry > P.hLy) ry -= P.Ly; else if (ry < -P.hLy) ry += P.Ly;
    }
}

// Stencil of particle p: z range clipped at the walls (asph_h100.cu:301-303) and the spread
// (Morton) bits of every wrapped x column.
template <int S> struct Stencil {
    int ix, iy, iz, z0, z1;
    uint32_t px[2 * S + 1];
    __device__ __forceinline__ Stencil(const DevParams& P, float x, float y, float z) {
        cell_trunc(P, x, y, z, ix, iy, iz);
        bool interior = (iz >= S && iz < P.nz - S);
        z0 = interior ? iz - S : max(iz - S, 0);
        z1 = interior ? iz + S : min(iz + S, P.nz - 1);
#pragma unroll
        for (int d = 0; d < 2 * S + 1; ++d) px[d] = part1by2((uint32_t)wrapi(ix - S + d, P.nx));
    }
};

// ---------------------------------------------------------------------------------------------
// kick(dt/2) + drift, asph_h100.cu:535-561. Also clears the per-step rain deposit grid
// (consumed by k_rain of the previous step).
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_kick_drift(DevParams P, float4* __restrict__ pos,
                                                    float4* __restrict__ vel,
                                                    const float4* __restrict__ acc,
                                                    int64_t* __restrict__ dep, float hdt, float dt) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.N) return;
    float4 v = vel[i];
    float4 a = acc[i];
    v.x = __builtin_fmaf(a.x, hdt, v.x);
    v.y = __builtin_fmaf(a.y, hdt, v.y);
    v.z = __builtin_fmaf(a.z, hdt, v.z);
    float4 p = pos[i];
    p.x = __builtin_fmaf(v.x, dt, p.x);
    p.y = __builtin_fmaf(v.y, dt, p.y);
    p.z = __builtin_fmaf(v.z, dt, p.z);
    // periodic wrap (the reference's while loops, bounded so a non-finite position cannot spin)
    for (int k = 0; k < 8 && p.x < 0.f; ++k) p.x += P.Lx;
    for (int k = 0; k < 8 && p.x >= P.Lx; ++k) p.x -= P.Lx;
    for (int k = 0; k < 8 && p.y < 0.f; ++k) p.y += P.Ly;
    for (int k = 0; k < 8 && p.y >= P.Ly; ++k) p.y -= P.Ly;
    if (p.z < 0.f) { p.z = -p.z; v.z = fabsf(v.z); }
    if (p.z > P.Lz) { p.z = 2.f * P.Lz - p.z; v.z = -fabsf(v.z); }
    pos[i] = p;
    vel[i] = v;
    if (i < P.ncell) dep[i] = 0;
}

// ---------------------------------------------------------------------------------------------
// Surface flux (asph_h100.cu:606-626) + lagged-Shepard density (:286-333) + EOS, saturation
// adjustment, autoconversion (:340-402) + Kessler accretion and rain evaporation.
// The flux is applied first: the density loop does not read thq, so the reference's order
// flux -> density -> eos is kept.
// ---------------------------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(TPB) void k_density_eos(
    DevParams P, const float4* __restrict__ pos, const float* __restrict__ rho_old,
    float* __restrict__ rho_new, float4* __restrict__ thq, float* __restrict__ prs,
    const uint32_t* __restrict__ cstart, const uint32_t* __restrict__ keys_sorted,
    const int64_t* __restrict__ rain, int64_t* __restrict__ dep, uint32_t* __restrict__ maxv,
    float dt, float auto_f, float Ts, float qs_sfc, int do_eos) {
    // do_eos: bit 0 = EOS + microphysics after the density, bit 1 = surface flux first
    int i = blockIdx.x * TPB + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < 3 * 64) maxv[threadIdx.x] = 0u;
    if (i >= P.N) return;
    const float4 pi = pos[i];
    Stencil<S> st(P, pi.x, pi.y, pi.z);

    float S1 = 0.f, S2 = 0.f;
    for (int jz = st.z0; jz <= st.z1; ++jz) {
        uint32_t kz = part1by2((uint32_t)jz) << 2;
        for (int dy = 0; dy < 2 * S + 1; ++dy) {
            uint32_t kyz = kz | (part1by2((uint32_t)wrapi(st.iy - S + dy, P.ny)) << 1);
#pragma unroll
            for (int dx = 0; dx < 2 * S + 1; ++dx) {
                uint32_t k = kyz | st.px[dx];
                uint32_t s = cstart[k], e = cstart[k + 1];
                for (uint32_t j = s; j < e; ++j) {
                    float4 pj = pos[j];
                    float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                    min_image(P, rx, ry);
                    float dxm = rx * P.inv_dx, dym = ry * P.inv_dy, dzm = rz * P.inv_dz;
                    float r2 = __builtin_fmaf(dzm, dzm, __builtin_fmaf(dym, dym, dxm * dxm));
                    if (r2 < P.lim) {
                        float W = fker(sqrtf(r2) * P.inv_h);
                        S1 += W;
                        S2 += W * (1.0f / fmaxf(rho_old[j], 1e-3f));
                    }
                }
            }
        }
    }
    const float z = pi.z;
    const float ex = exner_ref(P, z);
    const float pref = pres_of_exner(P, ex);
    const float rref = pref / ((R_D * theta_ref(P, z)) * ex);
    float rr = S2 > 1e-20f ? S1 / S2 : rref;
    float rnew = fminf(fmaxf(rr, 0.5f * rref), 2.0f * rref);
    rho_new[i] = rnew;
    if (!do_eos) return;

    float4 t = thq[i];
    if ((do_eos & 2) && (keys_sorted[i] & 0x24924924u) == 0u && z < P.dz) {
        float Pi = asphm::fpow(fmaxf(pref, 1.0f) / P0, P.rd_cp);
        float Ta = t.x * Pi;
        t.x = t.x + (((P.c_h * P.v_bg) * (Ts - Ta)) * dt) / P.dz;
        t.y = t.y + ((((P.beta_moist * P.c_e) * P.v_bg) * (qs_sfc - t.y)) * dt) / P.dz;
        t.x = fminf(fmaxf(t.x, 100.f), P.theta_cap);
        t.y = fminf(fmaxf(t.y, 0.f), P.qv_cap);
    }

    // EOS about the hydrostatic reference, asph_h100.cu:353-362
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
    if (qv > qs || qc > 0.f) {  // Newton saturation adjustment, :370-385
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
    if ((P.flags & ASPH_F_AUTOCONV) && qc > P.qc_auto) {  // :389-394
        float dqr = (qc - P.qc_auto) * auto_f;
        qc -= dqr;
        dfx += to_fx(dqr);
    }
    int cid = -1;
    if (P.flags & (ASPH_F_ACCRETION | ASPH_F_RAIN_EVAP)) {
        cid = cell_of(P, pi.x, pi.y, pi.z);
        float rf = rain_fx_to_float(rain[cid]);
        float qr = (rf * P.mm) / r;
        if ((P.flags & ASPH_F_ACCRETION) && qc > 0.f && qr > 0.f) {
            float qcn = qc / (1.0f + (P.k_accr * dt) * asphm::fpow(qr, 0.875f));
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
                float dqe = fminf(fminf(E * dt, qr), qs2 - qv);
                qv += dqe;
                th -= (L_V * dqe) / (CP * Pi);
                dfx -= to_fx(dqe);
            }
        }
    }
    if (dfx != 0) {
        if (cid < 0) cid = cell_of(P, pi.x, pi.y, pi.z);
        atomicAdd((unsigned long long*)&dep[cid], (unsigned long long)dfx);
    }
    th = fminf(fmaxf(th, 100.0f), P.theta_cap);
    qv = fminf(fmaxf(qv, 0.0f), P.qv_cap);
    qc = fmaxf(qc, 0.0f);
    thq[i] = make_float4(th, qv, qc, 0.f);
    prs[i] = ppr;
}

// ---------------------------------------------------------------------------------------------
// Momentum (asph_h100.cu:415-511). mode 0 (force evaluation, :919): write acc and dthdt.
// mode 1 (step): write acc, then theta update (:562-571), second half kick, sponge (:518-530)
// into vel_out, and the per-axis max |v| (:631-647) reduced per wavefront and merged with
// atomicMax over 64 slots per axis.
// ---------------------------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(TPB) void k_momentum_kdk(
    DevParams P, const float4* __restrict__ pos, const float4* __restrict__ vel,
    float4* __restrict__ thq, const float* __restrict__ rho, const float* __restrict__ prs,
    const uint32_t* __restrict__ cstart, float4* __restrict__ acc, float* __restrict__ dthdt,
    float4* __restrict__ vel_out, uint32_t* __restrict__ maxv, float dt, int mode) {
    int i = blockIdx.x * TPB + threadIdx.x;
    float mx = 0.f, my = 0.f, mz = 0.f;
    if (i < P.N) {
        const float4 pi = pos[i];
        const float4 vi = vel[i];
        const float4 ti = thq[i];
        Stencil<S> st(P, pi.x, pi.y, pi.z);
        const float rhoi = fmaxf(rho[i], 1e-3f);
        const float inv_rhoi = 1.0f / rhoi;
        const float pri = prs[i];
        float ax = 0.f, ay = 0.f, az = 0.f, avx = 0.f, avy = 0.f, avz = 0.f;
        for (int jz = st.z0; jz <= st.z1; ++jz) {
            uint32_t kz = part1by2((uint32_t)jz) << 2;
            for (int dy = 0; dy < 2 * S + 1; ++dy) {
                uint32_t kyz = kz | (part1by2((uint32_t)wrapi(st.iy - S + dy, P.ny)) << 1);
#pragma unroll
                for (int dx = 0; dx < 2 * S + 1; ++dx) {
                    uint32_t k = kyz | st.px[dx];
                    uint32_t s = cstart[k], e = cstart[k + 1];
                    for (uint32_t j = s; j < e; ++j) {
                        if ((int)j == i) continue;
                        float4 pj = pos[j];
                        float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                        min_image(P, rx, ry);
                        float dxt = rx * P.inv_dx, dyt = ry * P.inv_dy, dzt = rz * P.inv_dz;
                        float r2t = __builtin_fmaf(dzt, dzt, __builtin_fmaf(dyt, dyt, dxt * dxt));
                        if (r2t >= P.lim || r2t < 1e-12f) continue;
                        float q = sqrtf(r2t) * P.inv_h;
                        float kk = kgrad(P, q);
                        float gx = (kk * dxt) * P.inv_dx;
                        float gy = (kk * dyt) * P.inv_dy;
                        float gz = (kk * dzt) * P.inv_dz;
                        float rhoj = fmaxf(rho[j], 1e-3f);
                        float inv_rhoj = 1.0f / rhoj;
                        float prj = prs[j];
                        float dp = (prj - pri) * (inv_rhoi * inv_rhoj);
                        float4 vj = vel[j];
                        float vdx = vi.x - vj.x, vdy = vi.y - vj.y, vdz = vi.z - vj.z;
                        float vdotr = __builtin_fmaf(vdz, rz, __builtin_fmaf(vdy, ry, vdx * rx));
                        float visc = 0.f;
                        if (vdotr < 0.f) {
                            float rp2 = __builtin_fmaf(rz, rz, __builtin_fmaf(ry, ry, rx * rx));
                            float rbar = 0.5f * (rhoi + rhoj);
                            visc = (P.visc_c * vdotr) / (((rp2 + P.eta2) * (rbar * rbar)) * rhoj);
                        }
                        float br = dp + visc;
                        ax = __builtin_fmaf(-br, gx, ax);
                        ay = __builtin_fmaf(-br, gy, ay);
                        az = __builtin_fmaf(-br, gz, az);
                        avx = __builtin_fmaf(-visc, gx, avx);
                        avy = __builtin_fmaf(-visc, gy, avy);
                        avz = __builtin_fmaf(-visc, gz, avz);
                    }
                }
            }
        }
        ax *= P.mm; ay *= P.mm; az *= P.mm;
        avx *= P.mm; avy *= P.mm; avz *= P.mm;
        const float z = pi.z;
        const float ex = exner_ref(P, z);
        const float thr = theta_ref(P, z);
        float thv = ti.x * (1.0f + 0.608f * ti.y);
        float thvr = thr * (1.0f + 0.608f * qv_ref(P, z));
        az = az + GRAV * ((thv - thvr) / thvr - ti.z);
        acc[i] = make_float4(ax, ay, az, 0.f);
        float Pi = asphm::fpow(fmaxf(pres_of_exner(P, ex) + pri, 1.0f) / P0, P.rd_cp);
        float dth = -__builtin_fmaf(vi.z, avz, __builtin_fmaf(vi.y, avy, vi.x * avx)) / (CP * Pi);
        dth = fminf(fmaxf(dth, -P.dthdt_cap), P.dthdt_cap);
        if (mode == 0) {
            dthdt[i] = dth;
        } else {
            float4 t = ti;
            t.x = fminf(fmaxf(__builtin_fmaf(dth, dt, t.x), 100.0f), P.theta_cap);
            thq[i] = t;
            const float hdt = 0.5f * dt;
            float4 v = vi;
            v.x = __builtin_fmaf(ax, hdt, v.x);
            v.y = __builtin_fmaf(ay, hdt, v.y);
            v.z = __builtin_fmaf(az, hdt, v.z);
            if ((P.flags & ASPH_F_SPONGE) && z >= P.sponge_zs) {
                float tt = (z - P.sponge_zs) / P.sponge_depth;
                float sg = (P.sponge_sigma * tt) * tt;
                v.z *= 1.0f / (1.0f + sg * dt);
            }
            vel_out[i] = v;
            mx = fabsf(v.x); my = fabsf(v.y); mz = fabsf(v.z);
        }
    }
    if (mode != 0) {
        // wave64 max, then one atomic per wave into slot (block % 64) of each axis
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            mx = fmaxf(mx, __shfl_xor(mx, o));
            my = fmaxf(my, __shfl_xor(my, o));
            mz = fmaxf(mz, __shfl_xor(mz, o));
        }
        if ((threadIdx.x & 63) == 0) {
            int slot = blockIdx.x & 63;
            atomicMax(&maxv[slot], __float_as_uint(mx));
            atomicMax(&maxv[64 + slot], __float_as_uint(my));
            atomicMax(&maxv[128 + slot], __float_as_uint(mz));
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Rain sedimentation, asph_h100.cu:577-599, on the exact fixed-point grid: the amount leaving a
// cell is trunc(cur * frac) and the same integer enters the cell below, so
// sum(rain) + sum(precip) is conserved exactly. Deposits of this step (dep) are merged first;
// a negative cell (rain evaporation exceeding the cell's rain) is clipped to zero and the
// clipped amount is counted.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_rain(DevParams P, const int64_t* __restrict__ rain_in,
                                              const int64_t* __restrict__ dep,
                                              int64_t* __restrict__ rain_out,
                                              int64_t* __restrict__ precip,
                                              unsigned long long* __restrict__ clipped, float frac,
                                              int sediment) {
    int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= P.ncell) return;
    int64_t cur = rain_in[c] + dep[c];
    if (cur < 0) {
        atomicAdd(clipped, (unsigned long long)(-cur));
        cur = 0;
    }
    if (!sediment) {
        rain_out[c] = cur;
        return;
    }
    int64_t om = (int64_t)((double)cur * (double)frac);
    int64_t im = 0;
    if (c + P.ncol < P.ncell) {
        int64_t ca = rain_in[c + P.ncol] + dep[c + P.ncol];
        if (ca < 0) ca = 0;
        im = (int64_t)((double)ca * (double)frac);
    }
    rain_out[c] = cur - om + im;
    if (c < P.ncol) precip[c] += om;
}

// Merge deposits into the rain grid without advection (init-time force evaluation, :918).
__global__ __launch_bounds__(TPB) void k_rain_merge(int n, int64_t* __restrict__ rain,
                                                    int64_t* __restrict__ dep) {
    int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= n) return;
    rain[c] += dep[c];
    dep[c] = 0;
}

// Gather, asph_h100.cu:245-256 (+ acc when ASPH_F_GATHER_ACC).
__global__ __launch_bounds__(TPB) void k_gather(int N, const uint32_t* __restrict__ perm,
                                                const float4* __restrict__ pA, float4* __restrict__ pB,
                                                const float4* __restrict__ vA, float4* __restrict__ vB,
                                                const float4* __restrict__ tA, float4* __restrict__ tB,
                                                const float* __restrict__ rA, float* __restrict__ rB,
                                                const float4* __restrict__ aA, float4* __restrict__ aB) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= N) return;
    uint32_t s = perm[i];
    pB[i] = pA[s];
    vB[i] = vA[s];
    tB[i] = tA[s];
    rB[i] = rA[s];
    if (aA) aB[i] = aA[s];
}

// Cell start table from the sorted keys: cstart[k] = first i with key_i >= k.
__global__ __launch_bounds__(TPB) void k_cell_start(int N, uint32_t nkeys,
                                                    const uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ cstart) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i < N && i > 0) {
        uint32_t k = keys[i], kp = keys[i - 1];
        for (uint32_t q = kp + 1; q <= k; ++q) cstart[q] = (uint32_t)i;
    }
    // head [0, key_0] -> 0 and tail (key_{N-1}, nkeys] -> N, grid-strided
    uint32_t k0 = keys[0], kl = keys[N - 1];
    uint32_t stride = gridDim.x * TPB;
    for (uint32_t q = (uint32_t)(blockIdx.x * TPB + threadIdx.x); q <= k0; q += stride) cstart[q] = 0u;
    for (uint32_t q = kl + 1 + (uint32_t)(blockIdx.x * TPB + threadIdx.x); q <= nkeys; q += stride)
        cstart[q] = (uint32_t)N;
}

// kernel_clamp_vel, asph_h100.cu:669-675.
__global__ __launch_bounds__(TPB) void k_clamp_vel(int N, float4* __restrict__ vel, float vmax) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= N) return;
    float4 v = vel[i];
    float s2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (s2 > vmax * vmax) {
        float s = sqrtf(s2), f = vmax / s;
        v.x *= f; v.y *= f; v.z *= f;
        vel[i] = v;
    }
}

// kernel_audit, asph_h100.cu:650-668 (FP64 accumulation) + rain water.
__global__ __launch_bounds__(TPB) void k_audit(DevParams P, const float4* __restrict__ pos,
                                               const float4* __restrict__ vel,
                                               const float4* __restrict__ thq,
                                               const int64_t* __restrict__ rain,
                                               const int64_t* __restrict__ dep,
                                               const int64_t* __restrict__ precip,
                                               double* __restrict__ tot) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    int stride = gridDim.x * TPB;
    for (int i = blockIdx.x * TPB + threadIdx.x; i < P.N; i += stride) {
        float4 p = pos[i], v = vel[i], t = thq[i];
        float Pi = asphm::fpow(fmaxf(pres_ref(P, p.z), 1.f) / P0, P.rd_cp);
        s[0] += 0.5 * (double)(v.x * v.x + v.y * v.y + v.z * v.z);
        s[1] += (double)(GRAV * p.z);
        s[2] += (double)((CP * t.x) * Pi);
        s[3] += (double)(t.y + t.z);
        s[4] += (double)v.z;
    }
    for (int c = blockIdx.x * TPB + threadIdx.x; c < P.ncell; c += stride)
        s[5] += (double)(rain[c] + (dep ? dep[c] : 0));
    for (int c = blockIdx.x * TPB + threadIdx.x; c < P.ncol; c += stride) s[5] += (double)precip[c];
    __shared__ double red[6][TPB];
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 6; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) atomicAdd(&tot[threadIdx.x], red[threadIdx.x][0]);
}

// kernel_coarse_dump, asph_h100.cu:680-703 (cells visited in the reference's order).
__global__ __launch_bounds__(TPB) void k_coarse_dump(DevParams P, const uint32_t* __restrict__ cstart,
                                                     const float4* __restrict__ vel,
                                                     const float4* __restrict__ thq,
                                                     float* __restrict__ out, int sx, int sy, int sz,
                                                     int cx, int cy, int cz) {
    int idx = blockIdx.x * TPB + threadIdx.x;
    if (idx >= cx * cy * cz) return;
    int cxx = idx % cx, cyy = (idx / cx) % cy, czz = idx / (cx * cy);
    float a[6] = {0, 0, 0, 0, 0, 0};
    int n = 0;
    int ix1 = min((cxx + 1) * sx, P.nx), iy1 = min((cyy + 1) * sy, P.ny), iz1 = min((czz + 1) * sz, P.nz);
    for (int iz = czz * sz; iz < iz1; ++iz)
        for (int iy = cyy * sy; iy < iy1; ++iy)
            for (int ix = cxx * sx; ix < ix1; ++ix) {
                uint32_t k = morton3((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
                for (uint32_t j = cstart[k]; j < cstart[k + 1]; ++j) {
                    float4 t = thq[j], v = vel[j];
                    a[0] += t.x; a[1] += t.y; a[2] += t.z;
                    a[3] += v.x; a[4] += v.y; a[5] += v.z;
                    n++;
                }
            }
    float inv = n ? 1.f / (float)n : 0.f;
    for (int c = 0; c < 6; ++c) out[(size_t)idx * 6 + c] = a[c] * inv;
}

// kernel_column_diag, asph_h100.cu:704-723.
__global__ __launch_bounds__(TPB) void k_column_diag(DevParams P, const uint32_t* __restrict__ cstart,
                                                     const float4* __restrict__ thq,
                                                     const float4* __restrict__ vel,
                                                     float4* __restrict__ out) {
    int col = blockIdx.x * TPB + threadIdx.x;
    if (col >= P.ncol) return;
    int iy = col / P.nx, ix = col % P.nx;
    float th = 0.f, qv = 0.f, qlmax = 0.f, wmax = 0.f;
    int n = 0;
    for (int k = 0; k < P.nz; ++k) {
        uint32_t key = morton3((uint32_t)ix, (uint32_t)iy, (uint32_t)k);
        for (uint32_t j = cstart[key]; j < cstart[key + 1]; ++j) {
            float4 t = thq[j];
            th += t.x; qv += t.y;
            qlmax = fmaxf(qlmax, t.z);
            wmax = fmaxf(wmax, fabsf(vel[j].z));
            n++;
        }
    }
    if (n) { float inv = 1.f / (float)n; th *= inv; qv *= inv; }
    out[col] = make_float4(th, qv, qlmax, wmax);
}

inline int nblocks(int64_t n) { return (int)((n + TPB - 1) / TPB); }

}  // namespace

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
void launch_kick_drift(const DevParams& P, hipStream_t s, float4* pos, float4* vel,
                       const float4* acc, int64_t* dep, float dt) {
    k_kick_drift<<<nblocks(P.N), TPB, 0, s>>>(P, pos, vel, acc, dep, 0.5f * dt, dt);
}

void launch_density_eos(const DevParams& P, hipStream_t s, const float4* pos, const float* rho_old,
                        float* rho_new, float4* thq, float* prs, const uint32_t* cstart,
                        const uint32_t* keys_sorted, const int64_t* rain, int64_t* dep,
                        uint32_t* maxv, float dt, float auto_f, float Ts, float qs_sfc, int do_eos) {
    dim3 g(nblocks(P.N));
    switch (P.S) {
        case 1: k_density_eos<1><<<g, TPB, 0, s>>>(P, pos, rho_old, rho_new, thq, prs, cstart, keys_sorted, rain, dep, maxv, dt, auto_f, Ts, qs_sfc, do_eos); break;
        case 2: k_density_eos<2><<<g, TPB, 0, s>>>(P, pos, rho_old, rho_new, thq, prs, cstart, keys_sorted, rain, dep, maxv, dt, auto_f, Ts, qs_sfc, do_eos); break;
        case 3: k_density_eos<3><<<g, TPB, 0, s>>>(P, pos, rho_old, rho_new, thq, prs, cstart, keys_sorted, rain, dep, maxv, dt, auto_f, Ts, qs_sfc, do_eos); break;
        default: k_density_eos<4><<<g, TPB, 0, s>>>(P, pos, rho_old, rho_new, thq, prs, cstart, keys_sorted, rain, dep, maxv, dt, auto_f, Ts, qs_sfc, do_eos); break;
    }
}

void launch_momentum_kdk(const DevParams& P, hipStream_t s, const float4* pos, const float4* vel,
                         float4* thq, const float* rho, const float* prs, const uint32_t* cstart,
                         float4* acc, float* dthdt, float4* vel_out, uint32_t* maxv, float dt,
                         int mode) {
    dim3 g(nblocks(P.N));
    switch (P.S) {
        case 1: k_momentum_kdk<1><<<g, TPB, 0, s>>>(P, pos, vel, thq, rho, prs, cstart, acc, dthdt, vel_out, maxv, dt, mode); break;
        case 2: k_momentum_kdk<2><<<g, TPB, 0, s>>>(P, pos, vel, thq, rho, prs, cstart, acc, dthdt, vel_out, maxv, dt, mode); break;
        case 3: k_momentum_kdk<3><<<g, TPB, 0, s>>>(P, pos, vel, thq, rho, prs, cstart, acc, dthdt, vel_out, maxv, dt, mode); break;
        default: k_momentum_kdk<4><<<g, TPB, 0, s>>>(P, pos, vel, thq, rho, prs, cstart, acc, dthdt, vel_out, maxv, dt, mode); break;
    }
}

void launch_rain(const DevParams& P, hipStream_t s, const int64_t* rain_in, const int64_t* dep,
                 int64_t* rain_out, int64_t* precip, unsigned long long* clipped, float frac,
                 int sediment) {
    k_rain<<<nblocks(P.ncell), TPB, 0, s>>>(P, rain_in, dep, rain_out, precip, clipped, frac, sediment);
}

void launch_rain_merge(hipStream_t s, int n, int64_t* rain, int64_t* dep) {
    k_rain_merge<<<nblocks(n), TPB, 0, s>>>(n, rain, dep);
}

void launch_gather(hipStream_t s, int N, const uint32_t* perm, const float4* pA, float4* pB,
                   const float4* vA, float4* vB, const float4* tA, float4* tB, const float* rA,
                   float* rB, const float4* aA, float4* aB) {
    k_gather<<<nblocks(N), TPB, 0, s>>>(N, perm, pA, pB, vA, vB, tA, tB, rA, rB, aA, aB);
}

void launch_cell_start(hipStream_t s, int N, uint32_t nkeys, const uint32_t* keys, uint32_t* cstart) {
    k_cell_start<<<nblocks(N), TPB, 0, s>>>(N, nkeys, keys, cstart);
}

void launch_clamp_vel(hipStream_t s, int N, float4* vel, float vmax) {
    k_clamp_vel<<<nblocks(N), TPB, 0, s>>>(N, vel, vmax);
}

void launch_audit(const DevParams& P, hipStream_t s, const float4* pos, const float4* vel,
                  const float4* thq, const int64_t* rain, const int64_t* dep, const int64_t* precip,
                  double* tot) {
    k_audit<<<1024, TPB, 0, s>>>(P, pos, vel, thq, rain, dep, precip, tot);
}

void launch_coarse_dump(const DevParams& P, hipStream_t s, const uint32_t* cstart, const float4* vel,
                        const float4* thq, float* out, int sx, int sy, int sz, int cx, int cy, int cz) {
    k_coarse_dump<<<nblocks((int64_t)cx * cy * cz), TPB, 0, s>>>(P, cstart, vel, thq, out, sx, sy, sz, cx, cy, cz);
}

void launch_column_diag(const DevParams& P, hipStream_t s, const uint32_t* cstart, const float4* thq,
                        const float4* vel, float4* out) {
    k_column_diag<<<nblocks(P.ncol), TPB, 0, s>>>(P, cstart, thq, vel, out);
}

}  // namespace asph
