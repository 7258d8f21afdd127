// asph_adjoint.hip - the reverse sweep of one KDK step on gfx950 (adjoint of column-integrated
// rain with respect to the initial state and the autoconversion threshold).
//
// One step's reverse runs the forward stages backwards: rain sedimentation (transpose), the
// momentum epilogue (pointwise vector-Jacobian product), the momentum pairs, the EOS and
// microphysics (pointwise), the density ratio (pointwise) and the density pairs, then kick and
// drift. Forward values come from the step's tape (recorded by the forward kernels while the
// segment is recomputed from its checkpoint), so no stage recomputes a neighbour sum.
//
// Pair stages gather: the thread of particle i visits the particles j listed near i's listed cell
// (cells culled against the support with the step's drift bound) and adds the derivative terms of
// the pairs (i, j) of i's own sum (j listed in the stencil of i's current cell) and (j, i) of j's
// sum (i listed in the stencil of j's current cell), each with the forward's support test on the
// same arithmetic. The neighbour sets are those of the forward, including its stale cell lists
// between sorts, and no atomics are needed.
// Reductions into the rain grid and the autoconversion-threshold derivative are FP64.
#include "asph_adjoint.h"
#include "asph_kernels.h"

namespace asph {

namespace {

constexpr int TPB = 256;

inline int nblk(int64_t n) { return (int)((n + TPB - 1) / TPB); }

__device__ __forceinline__ void min_image_adj(const DevParams& P, float& rx, float& ry) {
    if (P.flags & ASPH_F_MIN_IMAGE) {
        if (rx > P.hLx) rx -= P.Lx; else if (rx < -P.hLx) rx += P.Lx;
        if (ry > P.hLy) ry -= P.Ly; else if (ry < -P.hLy) ry += P.Ly;
    }
}

// Stencil z range clipped at the walls, asph_h100.cu:301-303.
__device__ __forceinline__ void zrange_adj(const DevParams& P, int iz, int& z0, int& z1) {
    const int S = P.S;
    bool interior = (iz >= S && iz < P.nz - S);
    z0 = interior ? iz - S : max(iz - S, 0);
    z1 = interior ? iz + S : min(iz + S, P.nz - 1);
}

// Whether cell (cx, cy, cz) lies in the stencil of current cell (ix, iy, iz) (periodic x, y).
__device__ __forceinline__ bool in_stencil(const DevParams& P, int ix, int iy, int iz, int cx, int cy, int cz) {
    int z0, z1;
    zrange_adj(P, iz, z0, z1);
    if (cz < z0 || cz > z1) return false;
    int dx = wrapi(cx - ix, P.nx), dy = wrapi(cy - iy, P.ny);
    if (dx > P.nx / 2) dx -= P.nx;
    if (dy > P.ny / 2) dy -= P.ny;
    return abs(dx) <= P.S && abs(dy) <= P.S;
}

__device__ __forceinline__ void listed_cell(uint32_t key, int& cx, int& cy, int& cz) {
    cx = (int)compact1by2(key);
    cy = (int)compact1by2(key >> 1);
    cz = (int)compact1by2(key >> 2);
}

__device__ __forceinline__ void add_f64(double* sink, double v) {
    // block-reduced by the caller's wave reduction; one atomic per wavefront
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0 && v != 0.0) atomicAdd(sink, v);
}

// ---- rain sedimentation transpose (asph_h100.cu:577-599) ---------------------------------------
__global__ __launch_bounds__(TPB) void k_adj_rain(const DevParams* __restrict__ Pp, float frac, int sediment,
                                                  const double* __restrict__ rb_end, double* __restrict__ rb_cur) {
    const DevParams& P = *Pp;
    const int64_t c = (int64_t)blockIdx.x * TPB + threadIdx.x;
    if (c >= P.N) return;
    const int64_t plane = (int64_t)P.nx * P.ny;
    double v = rb_end[c];
    if (sediment) {
        v = (1.0 - (double)frac) * v;
        if (c >= plane) v += (double)frac * rb_end[c - plane];
    }
    rb_cur[c] = v;
}

// ---- momentum epilogue (pointwise) -------------------------------------------------------------
// Seeds of the pair stage (Ab, AVb: adjoints of the pressure/viscous and viscous-only pair sums)
// and the pointwise adjoints of the stage's inputs.
__global__ __launch_bounds__(TPB) void k_adj_mom_seed(const DevParams* __restrict__ Pp, AdjTape T, AdjWork W,
                                                      AdjState E, float dt, int mode) {
    const DevParams& P = *Pp;
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.N) return;
    const float4 x1 = T.x1[i], v1 = T.v1[i], t1 = T.t1[i], a1 = T.a1[i], av = T.av[i];
    // the pair part of the acceleration: the stage's output minus the buoyancy
    float in[adj::MOM_IN];
    in[9] = x1.z; in[10] = t1.x; in[11] = t1.y; in[12] = t1.z; in[13] = v1.w;
    {
        float thv = t1.x * (1.0f + 0.608f * t1.y);
        float thvr = theta_ref(P, x1.z) * (1.0f + 0.608f * qv_ref(P, x1.z));
        float buoy = GRAV * ((thv - thvr) / thvr - t1.z);
        in[0] = a1.x; in[1] = a1.y; in[2] = a1.z - buoy;
    }
    in[3] = av.x; in[4] = av.y; in[5] = av.z;
    in[6] = v1.x; in[7] = v1.y; in[8] = v1.z;
    const float4 ab = E.a[i], vb = E.v[i], tb = E.t[i];
    float ob[adj::MOM_OUT] = {ab.x, ab.y, ab.z, tb.x, vb.x, vb.y, vb.z};
    float ib[adj::MOM_IN];
    adj::mom_vjp<float>(P, dt, mode, in, ob, ib);
    const float mm = P.mm;
    W.Ab[i] = make_float4(ib[0] * mm, ib[1] * mm, ib[2] * mm, 0.f);
    W.AVb[i] = make_float4(ib[3] * mm, ib[4] * mm, ib[5] * mm, 0.f);
    const float4 xb = E.x[i];
    W.x1b[i] = make_float4(xb.x, xb.y, xb.z + ib[9], 0.f);
    W.v1b[i] = make_float4(ib[6], ib[7], ib[8], 0.f);
    W.t2b[i] = make_float4(ib[10], tb.y + ib[11], tb.z + ib[12], 0.f);
    W.pb[i] = ib[13];
    W.r1b[i] = E.rho[i];
}

// ---- the candidates of a particle's pair stages ------------------------------------------------------
// The pairs of i's reverse are (i, j) of i's own sum (j listed in the stencil of i's current cell)
// and (j, i) of j's sum (i listed in the stencil of j's current cell). A particle listed in cell c
// lies within [c - d, c + 1 + d] per axis (d: the step's drift bound since the sort, with its
// rounding margin), so its current cell is within c +- (1 + floor d), and both kinds of partner
// are listed within M = S + 1 + floor(d) cells of i's listed cell. The walk visits those cells,
// skips every cell whose box [c - d, c + 1 + d] (nearest periodic image) lies outside the support
// around i (lim widened by 1e-4 relative), and calls f(j, own) for each particle listed in the rest,
// own telling whether the cell is in the stencil of i's current cell.
struct Around {
    float xc, yc, zc;   // position [cells]
    int lx, ly, lz;     // listed cell
    int ix, iy, iz;     // current cell
};

__device__ __forceinline__ Around around(const DevParams& P, float4 x, uint32_t key) {
    Around a;
    a.xc = x.x * P.inv_dx;
    a.yc = x.y * P.inv_dy;
    a.zc = x.z * P.inv_dz;
    listed_cell(key, a.lx, a.ly, a.lz);
    cell_trunc(P, x.x, x.y, x.z, a.ix, a.iy, a.iz);
    return a;
}

// Distance [cells] from t (the offset of a point from a cell's centre) to the cell's box widened
// by d, over the nearest periodic image when n > 0.
__device__ __forceinline__ float box_gap(float t, float d, int n) {
    if (n > 0) {
        const float hn = 0.5f * (float)n;
        if (t > hn) t -= (float)n;
        else if (t < -hn) t += (float)n;
    }
    return fmaxf(0.f, fabsf(t) - 0.5f - d);
}

template <class F>
__device__ __forceinline__ void walk_partners(const DevParams& P, const uint32_t* __restrict__ cstart,
                                              const float* __restrict__ dl, const Around& a, F f) {
    const float limc = P.lim * 1.0001f;
    const int S = P.S;
    const float d0 = dl[0], d1 = dl[1], d2 = dl[2];
    const int Mx = S + 1 + (int)d0, My = S + 1 + (int)d1, Mz = S + 1 + (int)d2;
    const int nxs = min(2 * Mx + 1, P.nx), nys = min(2 * My + 1, P.ny);
    const int za = max(a.lz - Mz, 0), zb = min(a.lz + Mz, P.nz - 1);
    for (int cz = za; cz <= zb; ++cz) {
        const float ez = box_gap(a.zc - ((float)cz + 0.5f), d2, 0);
        const float ez2 = ez * ez;
        if (ez2 >= limc) continue;
        for (int oy = 0; oy < nys; ++oy) {
            const int cyu = a.ly - My + oy;
            const float ey = box_gap(a.yc - ((float)cyu + 0.5f), d1, P.ny);
            const float eyz = __builtin_fmaf(ey, ey, ez2);
            if (eyz >= limc) continue;
            const int cy = wrapi(cyu, P.ny);
            for (int ox = 0; ox < nxs; ++ox) {
                const int cxu = a.lx - Mx + ox;
                const float ex = box_gap(a.xc - ((float)cxu + 0.5f), d0, P.nx);
                if (__builtin_fmaf(ex, ex, eyz) >= limc) continue;
                const int cx = wrapi(cxu, P.nx);
                const bool own = in_stencil(P, a.ix, a.iy, a.iz, cx, cy, cz);
                const uint32_t key = morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
                const uint32_t j0 = cstart[key], j1 = cstart[key + 1];
                for (uint32_t j = j0; j < j1; ++j) f(j, own);
            }
        }
    }
}

// The support test shared by both orientations of a pair (the forward's arithmetic; negating r is
// exact, so both orientations square to the same value) and the reverse-role stencil test.
__device__ __forceinline__ bool in_support(const DevParams& P, float rx, float ry, float rz) {
    const float dxm = rx * P.inv_dx, dym = ry * P.inv_dy, dzm = rz * P.inv_dz;
    const float r2 = __builtin_fmaf(dzm, dzm, __builtin_fmaf(dym, dym, dxm * dxm));
    return r2 < P.lim;
}

__device__ __forceinline__ bool rev_role(const DevParams& P, float4 xj, const Around& a) {
    int jx, jy, jz;
    cell_trunc(P, xj.x, xj.y, xj.z, jx, jy, jz);
    return in_stencil(P, jx, jy, jz, a.lx, a.ly, a.lz);
}

// ---- momentum pairs (gather) -------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_adj_mom_pairs(const DevParams* __restrict__ Pp, AdjTape T, AdjWork W) {
    const DevParams& P = *Pp;
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.N) return;
    const float4 xi = T.x1[i], vi4 = T.v1[i];
    const float vi[3] = {vi4.x, vi4.y, vi4.z};
    const float rho_i_raw = T.r1[i];
    const float rho_i = fmaxf(rho_i_raw, 1e-3f), p_i = vi4.w;
    const float4 Ai4 = W.Ab[i], AVi4 = W.AVb[i];
    const float Ai[3] = {Ai4.x, Ai4.y, Ai4.z}, AVi[3] = {AVi4.x, AVi4.y, AVi4.z};
    const Around a = around(P, xi, T.keys[i]);
    float xb[3] = {0.f, 0.f, 0.f}, vb[3] = {0.f, 0.f, 0.f}, rb = 0.f, pb = 0.f;
    walk_partners(P, T.cstart, T.delta, a, [&](uint32_t j, bool own) {
        const float4 xj = T.x1[j];
        float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        min_image_adj(P, rx, ry);
        if (!in_support(P, rx, ry, rz)) return;
        const bool rev = rev_role(P, xj, a);
        if (!own && !rev) return;
        const float4 vj4 = T.v1[j];
        const float vj[3] = {vj4.x, vj4.y, vj4.z};
        const float rho_j = fmaxf(T.r1[j], 1e-3f), p_j = vj4.w;
        adj::PairBar<float> b;
        if (own && adj::mom_pair_vjp<float>(P, rx, ry, rz, vi, vj, rho_i, rho_j, p_i, p_j, Ai, AVi, b)) {
            for (int c = 0; c < 3; ++c) { xb[c] += b.r[c]; vb[c] += b.vd[c]; }
            rb += b.rho_o;
            pb += b.p_o;
        }
        if (rev) {
            const float4 Aj4 = W.Ab[j], AVj4 = W.AVb[j];
            const float Aj[3] = {Aj4.x, Aj4.y, Aj4.z}, AVj[3] = {AVj4.x, AVj4.y, AVj4.z};
            if (adj::mom_pair_vjp<float>(P, -rx, -ry, -rz, vj, vi, rho_j, rho_i, p_j, p_i, Aj, AVj, b)) {
                for (int c = 0; c < 3; ++c) { xb[c] -= b.r[c]; vb[c] -= b.vd[c]; }
                rb += b.rho_n;
                pb += b.p_n;
            }
        }
    });
    float4 x1b = W.x1b[i], v1b = W.v1b[i];
    W.x1b[i] = make_float4(x1b.x + xb[0], x1b.y + xb[1], x1b.z + xb[2], 0.f);
    W.v1b[i] = make_float4(v1b.x + vb[0], v1b.y + vb[1], v1b.z + vb[2], 0.f);
    if (rho_i_raw > 1e-3f) W.r1b[i] += rb;
    W.pb[i] += pb;
}

// ---- EOS + microphysics (pointwise) --------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_adj_eos(const DevParams* __restrict__ Pp, AdjTape T, AdjWork W,
                                                 AdjState Sst, const double* __restrict__ rb_cur,
                                                 double* __restrict__ qc0b, float dt, float Ts, float qs_sfc,
                                                 int flux_step) {
    const DevParams& P = *Pp;
    const int i = blockIdx.x * TPB + threadIdx.x;
    double q0 = 0.0;
    if (i < P.N) {
        const float4 t0 = T.t0[i], x1 = T.x1[i];
        const int cid = cell_of(P, x1.x, x1.y, x1.z);
        adj::EosConst<float> K;
        K.dt = dt;
        K.auto_f = auto_factor(P, dt);
        K.Ts = Ts;
        K.qs_sfc = qs_sfc;
        K.flux = flux_step && (P.flags & ASPH_F_SURFACE_FLUX) && (T.keys[i] & 0x24924924u) == 0u && x1.z < P.dz;
        const float in[adj::EOS_IN] = {t0.x, t0.y, t0.z, T.r1[i], x1.z, rain_fx_to_float(T.rain0[cid]), P.qc_auto};
        const float4 t2b = W.t2b[i];
        const float ob[adj::EOS_OUT] = {t2b.x, t2b.y, t2b.z, W.pb[i], (float)rb_cur[cid]};
        float ib[adj::EOS_IN];
        adj::eos_vjp<float>(P, K, in, ob, ib);
        Sst.t[i] = make_float4(ib[0], ib[1], ib[2], 0.f);
        W.r1b[i] += ib[3];
        W.x1b[i].z += ib[4];
        if (ib[5] != 0.f) atomicAdd(&Sst.rain[cid], (double)ib[5]);
        q0 = (double)ib[6];
    }
    add_f64(qc0b, q0);
}

// ---- density ratio (pointwise) ------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_adj_dens_seed(const DevParams* __restrict__ Pp, AdjTape T, AdjWork W) {
    const DevParams& P = *Pp;
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.N) return;
    const float2 s12 = T.s12[i];
    float s1b, s2b, zb;
    adj::dens_vjp<float>(P, s12.x, s12.y, T.x1[i].z, W.r1b[i], s1b, s2b, zb);
    W.sb[i] = make_float2(s1b, s2b);
    W.x1b[i].z += zb;
}

// ---- density pairs (gather) ---------------------------------------------------------------------
// W and its gradient are even and odd in r, so one evaluation serves both orientations.
__global__ __launch_bounds__(TPB) void k_adj_dens_pairs(const DevParams* __restrict__ Pp, AdjTape T, AdjWork W,
                                                        AdjState Sst) {
    const DevParams& P = *Pp;
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.N) return;
    const float4 xi = T.x1[i];
    const float r0i_raw = T.r0[i];
    const float inv_r0i = 1.0f / fmaxf(r0i_raw, 1e-3f);
    const float2 sbi = W.sb[i];
    const Around a = around(P, xi, T.keys[i]);
    float xb[3] = {0.f, 0.f, 0.f}, rb = 0.f;
    walk_partners(P, T.cstart, T.delta, a, [&](uint32_t j, bool own) {
        const float4 xj = T.x1[j];
        float rx = xi.x - xj.x, ry = xi.y - xj.y, rz = xi.z - xj.z;
        min_image_adj(P, rx, ry);
        if (!in_support(P, rx, ry, rz)) return;
        const bool rev = rev_role(P, xj, a);
        if (!own && !rev) return;
        float Wv, gw[3];
        if (!adj::dens_pair<float>(P, rx, ry, rz, Wv, gw)) return;
        float wb = 0.f;
        if (own) wb = sbi.x + sbi.y * (1.0f / fmaxf(T.r0[j], 1e-3f));
        if (rev) {
            // (j, i): r_ji = -r_ij, so the gradient flips sign and x_i enters with a minus
            const float2 sbj = W.sb[j];
            wb += sbj.x + sbj.y * inv_r0i;
            rb += sbj.y * Wv * (-inv_r0i * inv_r0i);
        }
        for (int c = 0; c < 3; ++c) xb[c] += wb * gw[c];
    });
    float4 x1b = W.x1b[i];
    W.x1b[i] = make_float4(x1b.x + xb[0], x1b.y + xb[1], x1b.z + xb[2], 0.f);
    Sst.rho[i] = r0i_raw > 1e-3f ? rb : 0.f;
}

// ---- kick + drift (asph_h100.cu:535-561) ------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_adj_kd(const DevParams* __restrict__ Pp, AdjTape T, AdjWork W, AdjState Sst,
                                                float dt, int mode) {
    const DevParams& P = *Pp;
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= P.N) return;
    const float4 xb = W.x1b[i], vb = W.v1b[i];
    if (mode == 0) {   // the force evaluation at t = 0 moves nothing
        Sst.x[i] = xb;
        Sst.v[i] = vb;
        Sst.a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const float4 x0 = T.x0[i], v0 = T.v0[i], a0 = T.a0[i];
    const float hdt = 0.5f * dt;
    const float vz = __builtin_fmaf(a0.z, hdt, v0.z);
    const float z1 = __builtin_fmaf(vz, dt, x0.z);
    const float s = (z1 < 0.f || z1 > P.Lz) ? -1.f : 1.f;   // wall reflection flips z and w
    const float zb = s * xb.z;
    const float vpx = vb.x + dt * xb.x, vpy = vb.y + dt * xb.y, vpz = s * vb.z + dt * zb;
    Sst.x[i] = make_float4(xb.x, xb.y, zb, 0.f);
    Sst.v[i] = make_float4(vpx, vpy, vpz, 0.f);
    Sst.a[i] = make_float4(hdt * vpx, hdt * vpy, hdt * vpz, 0.f);
}

// ---- the drift bound of a step's neighbour stages ------------------------------------------------
// The same fold as the forward kernels' prologue: the drift since the sort plus the step's
// per-slot maxima, plus the 1e-3 cell rounding margin of the culling bounds.
__global__ __launch_bounds__(64) void k_adj_delta(const float* __restrict__ disp, const uint32_t* __restrict__ slots,
                                                 float* __restrict__ out) {
    const int t = threadIdx.x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float v = __uint_as_float(slots[c * 64 + t]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        if (t == 0) out[c] = disp[c] + v + 1e-3f;
    }
}

// ---- undo a sort: adjoints of the gathered arrays scatter back through the permutation ----------
__global__ __launch_bounds__(TPB) void k_adj_unperm(int N, const uint32_t* __restrict__ perm, AdjState in,
                                                    AdjState out) {
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= N) return;
    const uint32_t k = perm[i];
    out.x[k] = in.x[i];
    out.v[k] = in.v[i];
    out.a[k] = in.a[i];
    out.t[k] = in.t[i];
    out.rho[k] = in.rho[i];
}

}  // namespace

void launch_adj_step(const DevParams* P, int64_t N, hipStream_t s, const AdjTape& T, const AdjWork& W,
                     const AdjState& E, const AdjState& Sst, double* qc0b, float dt, float Ts, float qs_sfc,
                     float frac, int mode) {
    const int g = nblk(N);
    // rain: sedimentation at a step, the plain merge of the deposits at the force evaluation
    k_adj_rain<<<g, TPB, 0, s>>>(P, frac, mode != 0 && frac > 0.f ? 1 : 0, E.rain, W.rb_cur);
    k_adj_mom_seed<<<g, TPB, 0, s>>>(P, T, W, E, dt, mode);
    k_adj_mom_pairs<<<g, TPB, 0, s>>>(P, T, W);
    // the rain grid at the start of the stage: what entered the sedimentation, plus the reads
    (void)hipMemcpyAsync(Sst.rain, W.rb_cur, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s);
    k_adj_eos<<<g, TPB, 0, s>>>(P, T, W, Sst, W.rb_cur, qc0b, dt, Ts, qs_sfc, mode);
    k_adj_dens_seed<<<g, TPB, 0, s>>>(P, T, W);
    k_adj_dens_pairs<<<g, TPB, 0, s>>>(P, T, W, Sst);
    k_adj_kd<<<g, TPB, 0, s>>>(P, T, W, Sst, dt, mode);
}

void launch_adj_delta(hipStream_t s, const float* disp, const uint32_t* slots, float* out) {
    k_adj_delta<<<1, 64, 0, s>>>(disp, slots, out);
}

void launch_adj_unperm(int64_t N, hipStream_t s, const uint32_t* perm, const AdjState& in, const AdjState& out) {
    k_adj_unperm<<<nblk(N), TPB, 0, s>>>((int)N, perm, in, out);
    // the rain grid is per cell: no permutation
    (void)hipMemcpyAsync(out.rain, in.rain, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s);
}

}  // namespace asph
