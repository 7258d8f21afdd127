// asph_adjoint.hip - the reverse sweep of one KDK step on gfx950 (adjoint of column-integrated
// rain with respect to the initial state and the autoconversion threshold).
//
// One step's reverse runs the forward stages backwards: rain sedimentation (transpose), the
// momentum epilogue (pointwise vector-Jacobian product), the momentum pairs, the EOS and
// microphysics (pointwise), the density ratio (pointwise) and the density pairs, then kick and
// drift. Forward values come from the step's tape (recorded by the forward kernels while the
// segment is recomputed from its checkpoint), so no stage recomputes a neighbour sum.
//
// The pair stages (asph_nbr.hip, launch_adj_pairs) gather: for particle i, the derivative terms of
// the pairs (i, j) of i's own sum and (j, i) of j's sums, on the forward's staged halo. The
// neighbour sets are those of the forward, including its stale cell lists between sorts, and no
// atomics are needed.
// Reductions into the rain grid and the autoconversion-threshold derivative are FP64.
#include <algorithm>

#include "asph_adjoint.h"
#include "asph_kernels.h"

namespace asph {

namespace {

constexpr int TPB = 256;

inline int nblk(int64_t n) { return (int)((n + TPB - 1) / TPB); }

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
    adj::mom_vjp_rev<float>(P, dt, mode, in, ob, ib);
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

// ---- EOS + microphysics (pointwise) --------------------------------------------------------------
// Evaluated in double at the forward's FP32 values and constants: the saturation adjustment's
// Newton iterations lose about 1e-4 of the derivative per step in float, which accumulates over a
// window; the other stages keep float.
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
        adj::EosConst<double> K;
        K.dt = dt;
        K.auto_f = auto_factor(P, dt);   // the forward's float constant
        K.Ts = Ts;
        K.qs_sfc = qs_sfc;
        K.flux = flux_step && (P.flags & ASPH_F_SURFACE_FLUX) && (T.keys[i] & 0x24924924u) == 0u && x1.z < P.dz;
        const double in[adj::EOS_IN] = {t0.x, t0.y, t0.z, T.r1[i], x1.z, rain_fx_to_float(T.rain0[cid]), P.qc_auto};
        const float4 t2b = W.t2b[i];
        const double ob[adj::EOS_OUT] = {t2b.x, t2b.y, t2b.z, W.pb[i], rb_cur[cid]};
        double ib[adj::EOS_IN];
        adj::eos_vjp<double>(P, K, in, ob, ib);
        Sst.t[i] = make_float4((float)ib[0], (float)ib[1], (float)ib[2], 0.f);
        W.r1b[i] += (float)ib[3];
        W.x1b[i].z += (float)ib[4];
        if (ib[5] != 0.0) atomicAdd(&Sst.rain[cid], ib[5]);
        q0 = ib[6];
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

// ---- tape records: several device copies in one launch (y: the entry) ---------------------------
__global__ __launch_bounds__(TPB) void k_adj_copy(AdjCopies C) {
    const int e = blockIdx.y;
    const uint32_t n = C.words[e], n4 = n / 4;
    const uint4* __restrict__ s4 = (const uint4*)C.src[e];
    uint4* __restrict__ d4 = (uint4*)C.dst[e];
    const uint32_t stride = gridDim.x * TPB;
    for (uint32_t k = blockIdx.x * TPB + threadIdx.x; k < n4; k += stride) d4[k] = s4[k];
    for (uint32_t k = 4 * n4 + blockIdx.x * TPB + threadIdx.x; k < n; k += stride) C.dst[e][k] = C.src[e][k];
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

void launch_adj_step(const Launch& L, int S, const DevParams* P, int64_t N, hipStream_t s, const AdjTape& T,
                     const AdjWork& W, const AdjState& E, const AdjState& Sst, double* qc0b, float dt, float Ts,
                     float qs_sfc, float frac, int mode) {
    const int g = nblk(N);
    // rain: sedimentation at a step, the plain merge of the deposits at the force evaluation
    k_adj_rain<<<g, TPB, 0, s>>>(P, frac, mode != 0 && frac > 0.f ? 1 : 0, E.rain, W.rb_cur);
    k_adj_mom_seed<<<g, TPB, 0, s>>>(P, T, W, E, dt, mode);
    launch_adj_pairs(L, S, s, 0, T, W, Sst);
    // the rain grid at the start of the stage: what entered the sedimentation, plus the reads
    (void)hipMemcpyAsync(Sst.rain, W.rb_cur, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s);
    k_adj_eos<<<g, TPB, 0, s>>>(P, T, W, Sst, W.rb_cur, qc0b, dt, Ts, qs_sfc, mode);
    k_adj_dens_seed<<<g, TPB, 0, s>>>(P, T, W);
    launch_adj_pairs(L, S, s, 1, T, W, Sst);
    k_adj_kd<<<g, TPB, 0, s>>>(P, T, W, Sst, dt, mode);
}

void launch_adj_delta(hipStream_t s, const float* disp, const uint32_t* slots, float* out) {
    k_adj_delta<<<1, 64, 0, s>>>(disp, slots, out);
}

void launch_adj_copies(hipStream_t s, const AdjCopies& C) {
    if (C.count <= 0) return;
    uint32_t mx = 0;
    for (int e = 0; e < C.count; ++e) mx = C.words[e] > mx ? C.words[e] : mx;
    const uint32_t g = std::min<uint32_t>((mx / 4 + TPB - 1) / TPB + 1, 2048u);
    k_adj_copy<<<dim3(g, (unsigned)C.count), TPB, 0, s>>>(C);
}

void launch_adj_unperm(int64_t N, hipStream_t s, const uint32_t* perm, const AdjState& in, const AdjState& out) {
    k_adj_unperm<<<nblk(N), TPB, 0, s>>>((int)N, perm, in, out);
    // the rain grid is per cell: no permutation
    (void)hipMemcpyAsync(out.rain, in.rain, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, s);
}

}  // namespace asph
