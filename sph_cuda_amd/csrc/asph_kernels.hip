// asph_kernels.hip - the streaming kernels of one KDK step (kick/drift, rain sedimentation),
// the sort helpers (Morton keys, gather, cell start table) and the diagnostics.
//
// Launch shape: 256 work-items (4 wave64) per workgroup. Every workgroup reads its member's
// parameters from the member table through the block map (one member per block), so one launch
// serves a single domain or a whole ensemble. Particles are stored in Morton order after every
// sort, so the 64 particles of a wavefront occupy a 4x4x4 block of cells.
//
// Neighbour structure: cstart[k] = index of the first particle whose key is >= k, for every key
// k in [0, nkeys], so cell k holds particles [cstart[k], cstart[k+1]) -- one 4-byte start table
// in key order instead of the reference's int2 cellRange in row-major order
// (asph_h100.cu:257-278), written with coalesced stores and no clearing pass.
//
// The two neighbour-loop kernels live in asph_nbr.hip.
#include "asph_kernels.h"

namespace asph {

namespace {

constexpr int TPB = 256;

// ---------------------------------------------------------------------------------------------
// kick(dt/2) + drift, asph_h100.cu:535-561. Also clears the per-step rain deposit grid
// (consumed by k_rain of the previous step) and records the step's drift bound.
//
// Grid-stride over the particle blocks with at most KD_GRID workgroups: each workgroup reduces
// its drift maxima in registers and LDS and issues three atomics at the end, so the 64 slots per
// axis see a few dozen atomics each (one atomic per wavefront serialised ~4k device-scope
// atomics on every slot at 256^3).
// ---------------------------------------------------------------------------------------------
constexpr int KD_GRID = 2048;

// How far a particle lies beyond its listed cell, in cells per axis: the distance of its scaled
// coordinate from the cell's interval [c, c + 1], the periodic axes by minimum image (the cell
// from its sort key, truncated and clamped as the key was).
__device__ __forceinline__ float excursion(float xs, int c, float period, float half, bool periodic) {
    float d = xs - ((float)c + 0.5f);
    if (periodic) d += fabsf(d) > half ? copysignf(period, -d) : 0.f;
    return fmaxf(fabsf(d) - 0.5f, 0.f);
}

__global__ __launch_bounds__(TPB) void k_kick_drift(const MemberDesc* __restrict__ mem,
                                                    const BlockMap* __restrict__ pmap, int npblk,
                                                    float4* __restrict__ pos, float4* __restrict__ vel,
                                                    const float4* __restrict__ acc,
                                                    int64_t* __restrict__ dep,
                                                    uint32_t* __restrict__ disp_slots, float hdt,
                                                    float dt, const StepCtl* __restrict__ ctl,
                                                    const uint32_t* __restrict__ keys,
                                                    float4* __restrict__ wexc) {
    __shared__ float red[3][TPB / 64];
    if (ctl) {
        if (ctl->idle) return;
        dt = ctl->dt;
        hdt = 0.5f * dt;
    }
    float mx = 0.f, my = 0.f, mz = 0.f;
    for (int b = blockIdx.x; b < npblk; b += gridDim.x) {   // uniform over the workgroup
        const BlockMap bm = pmap[b];
        const MemberDesc& M = mem[bm.m];
        const DevParams& P = M.P;
        const int il = bm.base + threadIdx.x;
        float ex = 0.f, ey = 0.f, ez = 0.f;
        if (il < P.N) {
            const int i = M.p0 + il;
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
            dep[i] = 0;   // cells = particles per member: cell index = particle index
            // drift bound of this step in cells per axis (|dx| <= |v dt|, also across a wrap or a wall)
            mx = fmaxf(mx, fabsf(v.x * dt) * P.inv_dx);
            my = fmaxf(my, fabsf(v.y * dt) * P.inv_dy);
            mz = fmaxf(mz, fabsf(v.z * dt) * P.inv_dz);
            if (keys) {
                // beyond the listed cell (the key of the last sort), on the scaled coordinates the
                // neighbour kernels difference; a non-finite position gives +inf or NaN, which no
                // bound test accepts
                const uint32_t kl = keys[i] & ((1u << P.end_bit) - 1u);
                ex = excursion(p.x * P.inv_dx, (int)compact1by2(kl), P.fnx, P.hnx, true);
                ey = excursion(p.y * P.inv_dy, (int)compact1by2(kl >> 1), P.fny, P.hny, true);
                ez = excursion(p.z * P.inv_dz, (int)compact1by2(kl >> 2), 0.f, 0.f, false);
                if (!(ex == ex && ey == ey && ez == ez)) ex = ey = ez = __int_as_float(0x7f800000);
            }
        }
        if (keys) {
            // wavefront maxima, one entry per wavefront of the block (member-local index / 64)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                ex = fmaxf(ex, __shfl_xor(ex, o));
                ey = fmaxf(ey, __shfl_xor(ey, o));
                ez = fmaxf(ez, __shfl_xor(ez, o));
            }
            if ((threadIdx.x & 63) == 0 && il < P.N)
                wexc[(size_t)b * (TPB / 64) + (threadIdx.x >> 6)] = make_float4(ex, ey, ez, 0.f);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, o));
        my = fmaxf(my, __shfl_xor(my, o));
        mz = fmaxf(mz, __shfl_xor(mz, o));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = mx; red[1][w] = my; red[2][w] = mz; }
    __syncthreads();
    if (threadIdx.x < 3) {
        float m = red[threadIdx.x][0];
#pragma unroll
        for (int k = 1; k < TPB / 64; ++k) m = fmaxf(m, red[threadIdx.x][k]);
        // non-negative floats order like their bit patterns
        atomicMax(&disp_slots[threadIdx.x * 64 + (blockIdx.x & 63)], __float_as_uint(m));
    }
}

// Per-brick drift bound, one wavefront per brick of the brick map: the largest excursion beyond
// their listed cells of the particles listed in the brick (in key order one index range,
// [cstart[key], cstart[key + 512])), from the wavefront maxima of the kick/drift kernel. An entry
// covering particles of a neighbouring brick as well only raises the bound.
__global__ __launch_bounds__(TPB) void k_brick_exc(const MemberDesc* __restrict__ mem,
                                                   const BlockMap* __restrict__ bmap, int nbblk,
                                                   const uint32_t* __restrict__ cstart,
                                                   const float4* __restrict__ wexc, float4* __restrict__ bexc,
                                                   const StepCtl* __restrict__ ctl) {
    if (ctl && ctl->idle) return;
    const int g = blockIdx.x * (TPB / 64) + (threadIdx.x >> 6);   // uniform over the wavefront
    if (g >= nbblk) return;
    const int lane = threadIdx.x & 63;
    const BlockMap bm = bmap[g];
    const MemberDesc& M = mem[bm.m];
    const uint32_t key = M.k0 + (uint32_t)bm.base;
    // the member's index range bounds the reads whatever the table holds
    const uint32_t p0 = (uint32_t)M.p0, p1 = p0 + (uint32_t)M.P.N;
    const uint32_t lo = min(max(cstart[key], p0), p1), hi = min(max(cstart[key + (1u << BRICK_BITS)], p0), p1);
    float mx = 0.f, my = 0.f, mz = 0.f;
    if (hi > lo) {
        const float4* W = wexc + (size_t)M.first_block * (TPB / 64);
        const uint32_t e1 = (hi - 1u - p0) >> 6;
        for (uint32_t e = ((lo - p0) >> 6) + (uint32_t)lane; e <= e1; e += 64) {
            const float4 v = W[e];
            mx = fmaxf(mx, v.x);
            my = fmaxf(my, v.y);
            mz = fmaxf(mz, v.z);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        mx = fmaxf(mx, __shfl_xor(mx, o));
        my = fmaxf(my, __shfl_xor(my, o));
        mz = fmaxf(mz, __shfl_xor(mz, o));
    }
    if (lane == 0) bexc[key >> BRICK_BITS] = make_float4(mx, my, mz, 0.f);
}

// Per-step bookkeeping, one wavefront: folds the drift maxima of the last step into the bound
// since the sort (reset != 0 after a sort clears it) and sets or advances the simulated-time
// clock read by the kernels (mode 0: clock = t, mode 1: clock += dt, mode 2: leave).
__global__ void k_step_begin(float* __restrict__ disp, uint32_t* __restrict__ slots, int reset,
                             double* __restrict__ clock, int mode, double t, float dt,
                             const StepCtl* __restrict__ ctl, int sort_gate) {
    if (ctl && (ctl->idle || (sort_gate && !ctl->do_sort))) return;
    int tid = threadIdx.x;   // 64 threads
    float m[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v = __uint_as_float(slots[a * 64 + tid]);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        m[a] = v;
        slots[a * 64 + tid] = 0u;
    }
    if (tid < 3) disp[tid] = reset ? 0.f : disp[tid] + m[tid];
    if (tid == 0) {
        if (mode == 0) clock[0] = t;
        else if (mode == 1) clock[0] = clock[0] + (double)dt;
    }
}

// ---------------------------------------------------------------------------------------------
// Rain sedimentation, asph_h100.cu:577-599, on the exact fixed-point grid: the amount leaving a
// cell is trunc(cur * frac) and the same integer enters the cell below, so
// sum(rain) + sum(precip) is conserved exactly. Deposits of this step (dep) are merged first;
// a negative cell (rain evaporation exceeding the cell's rain) is clipped to zero and the
// clipped amount is counted.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_rain(const MemberDesc* __restrict__ mem,
                                              const BlockMap* __restrict__ pmap,
                                              const int64_t* __restrict__ rain_in,
                                              const int64_t* __restrict__ dep,
                                              int64_t* __restrict__ rain_out,
                                              int64_t* __restrict__ precip,
                                              unsigned long long* __restrict__ clipped, float dt,
                                              const StepCtl* __restrict__ ctl) {
    const BlockMap bm = pmap[blockIdx.x];
    const MemberDesc& M = mem[bm.m];
    const DevParams& P = M.P;
    const int cl = bm.base + threadIdx.x;
    if (cl >= P.ncell) return;
    const int c = M.p0 + cl;
    if (ctl) {
        if (ctl->idle) {   // an idle step still rotates the grid buffers
            rain_out[c] = rain_in[c];
            return;
        }
        dt = ctl->dt;
    }
    int64_t cur = rain_in[c] + dep[c];
    if (cur < 0) {
        atomicAdd(clipped, (unsigned long long)(-cur));
        cur = 0;
    }
    if (!(P.flags & ASPH_F_SEDIMENT)) {
        rain_out[c] = cur;
        return;
    }
    const float frac = rain_frac(P, dt);
    int64_t om = (int64_t)((double)cur * (double)frac);
    int64_t im = 0;
    if (cl + P.ncol < P.ncell) {
        int64_t ca = rain_in[c + P.ncol] + dep[c + P.ncol];
        if (ca < 0) ca = 0;
        im = (int64_t)((double)ca * (double)frac);
    }
    rain_out[c] = cur - om + im;
    if (cl < P.ncol) precip[M.col0 + cl] += om;
}

// Merge deposits into the rain grid without advection (init-time force evaluation, :918).
__global__ __launch_bounds__(TPB) void k_rain_merge(int n, int64_t* __restrict__ rain,
                                                    int64_t* __restrict__ dep) {
    int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= n) return;
    rain[c] += dep[c];
    dep[c] = 0;
}

// Morton keys of the current positions (kernel_morton_keys, asph_h100.cu:229-240), prefixed
// with the member's key offset.
__global__ __launch_bounds__(TPB) void k_keys(const MemberDesc* __restrict__ mem,
                                              const BlockMap* __restrict__ pmap,
                                              const float4* __restrict__ pos,
                                              uint32_t* __restrict__ keys, uint32_t* __restrict__ ngaps,
                                              const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) *ngaps = 0u;   // gap list of this sort's cell table
    const BlockMap bm = pmap[blockIdx.x];
    const MemberDesc& M = mem[bm.m];
    const DevParams& P = M.P;
    const int il = bm.base + threadIdx.x;
    if (il >= P.N) return;
    const int i = M.p0 + il;
    float4 p = pos[i];
    int ix, iy, iz;
    cell_trunc(P, p.x, p.y, p.z, ix, iy, iz);
    keys[i] = M.k0 | morton3((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
}

// Gather, asph_h100.cu:245-256 (+ acc when ASPH_F_GATHER_ACC).
__global__ __launch_bounds__(TPB) void k_gather(int N, const uint32_t* __restrict__ perm,
                                                const float4* __restrict__ pA, float4* __restrict__ pB,
                                                const float4* __restrict__ vA, float4* __restrict__ vB,
                                                const float4* __restrict__ tA, float4* __restrict__ tB,
                                                const float* __restrict__ rA, float* __restrict__ rB,
                                                const float4* __restrict__ aA, float4* __restrict__ aB,
                                                const StepCtl* __restrict__ ctl) {
    int i = blockIdx.x * TPB + threadIdx.x;
    if (i >= N) return;
    // a step without a sort (device controller) copies in place order: the buffers still rotate
    uint32_t s = (ctl && !ctl->do_sort) ? (uint32_t)i : perm[i];
    pB[i] = pA[s];
    vB[i] = vA[s];
    tB[i] = tA[s];
    rB[i] = rA[s];
    if (aA) aB[i] = aA[s];
}

// Cell start table from the sorted keys: cstart[k] = first i with key_i >= k.
// Cell start table from the sorted keys: cstart[q] = index of the first particle with key >= q.
// The particle that starts a run of equal keys (and the end sentinel) owns the run of empty keys
// before it: runs of up to GAP_SERIAL keys it fills itself; longer ones (sparse key spaces:
// ensemble members, boxes that are not powers of two) go to a list that whole workgroups fill.
constexpr uint32_t GAP_SERIAL = 256;

__global__ __launch_bounds__(TPB) void k_cell_start(int N, uint32_t nkeys,
                                                    const uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ cstart, uint4* __restrict__ gaps,
                                                    uint32_t* __restrict__ ngaps, uint32_t gapcap,
                                                    const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    const int i = blockIdx.x * TPB + threadIdx.x;
    if (i > N) return;
    uint32_t lo, hi;
    if (i == 0) { lo = 0u; hi = keys[0]; }
    else if (i == N) { lo = keys[N - 1] + 1u; hi = nkeys; }
    else {
        const uint32_t k = keys[i], kp = keys[i - 1];
        if (k == kp) return;
        lo = kp + 1u; hi = k;
    }
    if (lo > hi) return;
    if (hi - lo < GAP_SERIAL) {
        for (uint32_t q = lo; q <= hi; ++q) cstart[q] = (uint32_t)i;
        return;
    }
    const uint32_t g = atomicAdd(ngaps, 1u);
    if (g < gapcap) gaps[g] = make_uint4(lo, hi, (uint32_t)i, 0u);
    else for (uint32_t q = lo; q <= hi; ++q) cstart[q] = (uint32_t)i;   // not reached: gapcap bounds the runs
}

__global__ __launch_bounds__(TPB) void k_cell_gaps(const uint4* __restrict__ gaps, const uint32_t* __restrict__ ngaps,
                                                   uint32_t gapcap, uint32_t* __restrict__ cstart,
                                                   const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    const uint32_t n = min(*ngaps, gapcap);
    for (uint32_t g = blockIdx.x; g < n; g += gridDim.x) {
        const uint4 r = gaps[g];
        for (uint32_t q = r.x + threadIdx.x; q <= r.y; q += TPB) cstart[q] = r.z;
    }
}

// kernel_clamp_vel, asph_h100.cu:669-675: |v| <= vmax_abs, or <= cs_frac * cs of each member.
__global__ __launch_bounds__(TPB) void k_clamp_vel(const MemberDesc* __restrict__ mem,
                                                   const BlockMap* __restrict__ pmap,
                                                   float4* __restrict__ vel, float vmax_abs,
                                                   float cs_frac, const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->clamp) return;
    const BlockMap bm = pmap[blockIdx.x];
    const MemberDesc& M = mem[bm.m];
    const int il = bm.base + threadIdx.x;
    if (il >= M.P.N) return;
    const int i = M.p0 + il;
    const float vmax = vmax_abs > 0.f ? vmax_abs : cs_frac * M.P.cs;
    float4 v = vel[i];
    float s2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (s2 > vmax * vmax) {
        float s = sqrtf(s2), f = vmax / s;
        v.x *= f; v.y *= f; v.z *= f;
        vel[i] = v;
    }
}

// kernel_audit, asph_h100.cu:650-668 (FP64 accumulation, per member scaled by its particle
// mass) + rain water. tot[0..6] as asph_audit.
__global__ __launch_bounds__(TPB) void k_audit(const MemberDesc* __restrict__ mem,
                                               const BlockMap* __restrict__ pmap,
                                               const float4* __restrict__ pos,
                                               const float4* __restrict__ vel,
                                               const float4* __restrict__ thq,
                                               const int64_t* __restrict__ rain,
                                               const int64_t* __restrict__ dep,
                                               const int64_t* __restrict__ precip,
                                               double* __restrict__ tot) {
    const BlockMap bm = pmap[blockIdx.x];
    const MemberDesc& M = mem[bm.m];
    const DevParams& P = M.P;
    const int il = bm.base + threadIdx.x;
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    if (il < P.N) {
        const int i = M.p0 + il;
        float4 p = pos[i], v = vel[i], t = thq[i];
        float Pi = asphm::fpow(fmaxf(pres_ref(P, p.z), 1.f) / P0, P.rd_cp);
        const double m = (double)P.pmass;
        s[0] = m * (0.5 * (double)(v.x * v.x + v.y * v.y + v.z * v.z));
        s[1] = m * (double)(GRAV * p.z);
        s[2] = m * (double)((CP * t.x) * Pi);
        s[3] = m * (double)(t.y + t.z);
        s[4] = m * (double)v.z;
        s[5] = m;
        double r = (double)(rain[i] + (dep ? dep[i] : 0));
        if (il < P.ncol) r += (double)precip[M.col0 + il];
        s[6] = m * r * FX_INV;
    }
    __shared__ double red[7][TPB];
#pragma unroll
    for (int k = 0; k < 7; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 7; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 7) atomicAdd(&tot[threadIdx.x], red[threadIdx.x][0]);
}

// Coarse dump (kernel_coarse_dump, asph_h100.cu:680-703) for member m: one thread per block of
// sx x sy x sz cells (blocks x fastest, then y, then z; the high-edge blocks are partial). The six
// fields [theta, qv, qc, u, v, w] are summed over the block's cells in x-fastest order and over the
// particles listed in each cell in listed order, the reference's order, on which the dump's bytes
// depend; an empty block writes zeros.
__global__ __launch_bounds__(TPB) void k_coarse_dump(const MemberDesc* __restrict__ mem, int m,
                                                     const uint32_t* __restrict__ cstart,
                                                     const float4* __restrict__ vel,
                                                     const float4* __restrict__ thq,
                                                     float* __restrict__ out, int sx, int sy, int sz,
                                                     int nbx, int nby, int nbz) {
    const MemberDesc& M = mem[m];
    const DevParams& P = M.P;
    const int blk = blockIdx.x * TPB + threadIdx.x;
    if (blk >= nbx * nby * nbz) return;
    const int x0 = (blk % nbx) * sx, y0 = ((blk / nbx) % nby) * sy, z0 = (blk / (nbx * nby)) * sz;
    const int x1 = min(x0 + sx, P.nx), y1 = min(y0 + sy, P.ny), z1 = min(z0 + sz, P.nz);
    float sum[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int count = 0;
    for (int z = z0; z < z1; ++z)
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x) {
                const uint32_t key = M.k0 | morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
                for (uint32_t j = cstart[key], e = cstart[key + 1]; j < e; ++j) {
                    const float4 tq = thq[j], u = vel[j];
                    sum[0] += tq.x; sum[1] += tq.y; sum[2] += tq.z;
                    sum[3] += u.x; sum[4] += u.y; sum[5] += u.z;
                    ++count;
                }
            }
    const float w = count ? 1.f / (float)count : 0.f;
    float* dst = out + (size_t)blk * 6;
    for (int f = 0; f < 6; ++f) dst[f] = sum[f] * w;
}

// Column diagnostics (kernel_column_diag, asph_h100.cu:704-723) for member m: per column, the mean
// theta and qv and the maximum qc and |w| over the particles listed in its cells, bottom to top.
__global__ __launch_bounds__(TPB) void k_column_diag(const MemberDesc* __restrict__ mem, int m,
                                                     const uint32_t* __restrict__ cstart,
                                                     const float4* __restrict__ thq,
                                                     const float4* __restrict__ vel,
                                                     float4* __restrict__ out) {
    const MemberDesc& M = mem[m];
    const DevParams& P = M.P;
    const int c = blockIdx.x * TPB + threadIdx.x;
    if (c >= P.ncol) return;
    const uint32_t kxy = M.k0 | morton3((uint32_t)(c % P.nx), (uint32_t)(c / P.nx), 0u);
    float th_sum = 0.f, qv_sum = 0.f, qc_max = 0.f, w_max = 0.f;
    int count = 0;
    for (int z = 0; z < P.nz; ++z) {
        const uint32_t key = kxy | (part1by2((uint32_t)z) << 2);
        for (uint32_t j = cstart[key], e = cstart[key + 1]; j < e; ++j) {
            const float4 tq = thq[j];
            th_sum += tq.x;
            qv_sum += tq.y;
            qc_max = fmaxf(qc_max, tq.z);
            w_max = fmaxf(w_max, fabsf(vel[j].z));
            ++count;
        }
    }
    if (count) {
        const float w = 1.f / (float)count;
        th_sum *= w;
        qv_sum *= w;
    }
    out[c] = make_float4(th_sum, qv_sum, qc_max, w_max);
}

inline int nblocks(int64_t n) { return (int)((n + TPB - 1) / TPB); }

}  // namespace

// ---------------------------------------------------------------------------------------------
// Device-resident step controller: the host loop of asph_run (asph_h100.cu:958-1078) as two
// one-wave kernels around each captured step. Same float arithmetic as the host (the expressions
// are written in the same order), so a run under it is bit for bit asph_run's.
// ---------------------------------------------------------------------------------------------
__global__ void k_ctl_pre(StepCtl* __restrict__ c) {
    if (threadIdx.x != 0) return;
    c->clamp = 0;
    const bool stop = c->abort || c->pause || (c->sim_end > 0.0 && c->simt >= c->sim_end) ||
                      (c->max_steps > 0 && c->steps >= c->max_steps);
    if (stop) {
        c->idle = 1;
        c->do_sort = 0;
        return;
    }
    c->idle = 0;
    c->steps += 1;
    const bool sort = c->every || c->since_sort >= (double)c->sort_dt;   // :963
    c->do_sort = sort ? 1 : 0;
    if (sort) {
        c->since_sort = 0.0;
        c->sorts += 1;
    }
}

__global__ void k_ctl_post(StepCtl* __restrict__ c, const uint32_t* __restrict__ maxv) {
    if (c->idle) return;
    const int t = threadIdx.x;   // 64 threads: max|v| of the step from the momentum kernel's slots
    float u[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uint32_t v = maxv[a * 64 + t];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o));
        u[a] = __uint_as_float(v);
    }
    if (t != 0) return;
    const float dt = c->dt;
    c->simt += (double)dt;                                                             // :1038
    c->since_sort += (double)dt;
    c->u[0] = u[0]; c->u[1] = u[1]; c->u[2] = u[2];
    const float dreq = c->cfl * fminf(fminf(c->dx / (c->cs + u[0]), c->dy / (c->cs + u[1])),
                                      c->dz / (c->cs + u[2])) * 0.90f;                 // :1048-1049
    c->dt_required = dreq;
    float adopt = c->fixed ? dt : fminf(c->dt_cap, dreq);
    float dtn = dt;
    if (!c->fixed) {
        if (adopt < c->dt_floor) {                                                     // :1054-1071
            c->trips += 1;
            if (c->strict || c->trips > c->max_trip) {
                c->abort = 1;
                c->idle = 1;
                return;
            }
            c->clamp = 1;
            adopt = c->clamp_dt;
        } else if (adopt < 0.98f * dt) {
            c->trips = 0;                                                              // :1072
        }
        dtn = adopt;
    }
    c->dt = dtn;
    const float dz_move = 0.3f * c->dz / fmaxf(u[2], 0.5f);                            // :1075-1077
    const float dh_move = 0.3f * c->dx / fmaxf(fmaxf(u[0], u[1]), 0.5f);
    c->sort_dt = fminf(fmaxf(fminf(dz_move, dh_move), dtn), 15.0f);
    if (c->dump_every > 0.0 && c->simt >= c->next_dump) c->pause = 1;                  // :1081
}

int particle_block_size() { return TPB; }

// ---------------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------------
void launch_step_begin(hipStream_t s, float* disp, uint32_t* disp_slots, int reset, double* clock,
                       int clock_mode, double t, float dt, const StepCtl* ctl, int sort_gate) {
    k_step_begin<<<1, 64, 0, s>>>(disp, disp_slots, reset, clock, clock_mode, t, dt, ctl, sort_gate);
}

void launch_kick_drift(const Launch& L, hipStream_t s, float4* pos, float4* vel, const float4* acc,
                       int64_t* dep, uint32_t* disp_slots, float dt, const uint32_t* keys,
                       const uint32_t* cstart) {
    const int grid = L.npblk < KD_GRID ? L.npblk : KD_GRID;
    const bool per_brick = keys && cstart && L.wexc && L.bexc;
    k_kick_drift<<<grid, TPB, 0, s>>>(L.mem, L.pmap, L.npblk, pos, vel, acc, dep, disp_slots,
                                      0.5f * dt, dt, L.ctl, per_brick ? keys : nullptr, L.wexc);
    if (per_brick)
        k_brick_exc<<<(L.nbblk + TPB / 64 - 1) / (TPB / 64), TPB, 0, s>>>(L.mem, L.bmap, L.nbblk, cstart, L.wexc,
                                                                          L.bexc, L.ctl);
}

void launch_rain(const Launch& L, hipStream_t s, const int64_t* rain_in, const int64_t* dep,
                 int64_t* rain_out, int64_t* precip, unsigned long long* clipped, float dt) {
    k_rain<<<L.npblk, TPB, 0, s>>>(L.mem, L.pmap, rain_in, dep, rain_out, precip, clipped, dt, L.ctl);
}

void launch_rain_merge(hipStream_t s, int n, int64_t* rain, int64_t* dep) {
    k_rain_merge<<<nblocks(n), TPB, 0, s>>>(n, rain, dep);
}

void launch_keys(const Launch& L, hipStream_t s, const float4* pos, uint32_t* keys) {
    k_keys<<<L.npblk, TPB, 0, s>>>(L.mem, L.pmap, pos, keys, L.ngaps, L.ctl);
}

void launch_gather(hipStream_t s, int N, const uint32_t* perm, const float4* pA, float4* pB,
                   const float4* vA, float4* vB, const float4* tA, float4* tB, const float* rA,
                   float* rB, const float4* aA, float4* aB, const StepCtl* ctl) {
    k_gather<<<nblocks(N), TPB, 0, s>>>(N, perm, pA, pB, vA, vB, tA, tB, rA, rB, aA, aB, ctl);
}

// Disjoint runs longer than GAP_SERIAL keys: at most (nkeys + 1) / (GAP_SERIAL + 1) + 1 of them.
uint32_t cell_gap_capacity(uint32_t nkeys) { return (nkeys + 1u) / (GAP_SERIAL + 1u) + 2u; }

void launch_cell_start(const Launch& L, hipStream_t s, int N, uint32_t nkeys, const uint32_t* keys,
                       uint32_t* cstart) {
    k_cell_start<<<nblocks((int64_t)N + 1), TPB, 0, s>>>(N, nkeys, keys, cstart, L.gaps, L.ngaps, L.gapcap,
                                                         L.ctl);
    k_cell_gaps<<<1024, TPB, 0, s>>>(L.gaps, L.ngaps, L.gapcap, cstart, L.ctl);
}

void launch_clamp_vel(const Launch& L, hipStream_t s, float4* vel, float vmax, float cs_frac) {
    k_clamp_vel<<<L.npblk, TPB, 0, s>>>(L.mem, L.pmap, vel, vmax, cs_frac, nullptr);
}

void launch_clamp_vel_ctl(const Launch& L, hipStream_t s, float4* vel) {
    k_clamp_vel<<<L.npblk, TPB, 0, s>>>(L.mem, L.pmap, vel, 0.f, 0.5f, L.ctl);
}

void launch_ctl_pre(hipStream_t s, StepCtl* ctl) { k_ctl_pre<<<1, 64, 0, s>>>(ctl); }

void launch_ctl_post(hipStream_t s, StepCtl* ctl, const uint32_t* maxv) { k_ctl_post<<<1, 64, 0, s>>>(ctl, maxv); }

void launch_audit(const Launch& L, hipStream_t s, const float4* pos, const float4* vel,
                  const float4* thq, const int64_t* rain, const int64_t* dep, const int64_t* precip,
                  double* tot) {
    k_audit<<<L.npblk, TPB, 0, s>>>(L.mem, L.pmap, pos, vel, thq, rain, dep, precip, tot);
}

void launch_coarse_dump(const Launch& L, int m, hipStream_t s, const uint32_t* cstart,
                        const float4* vel, const float4* thq, float* out, int sx, int sy, int sz,
                        int cx, int cy, int cz) {
    k_coarse_dump<<<nblocks((int64_t)cx * cy * cz), TPB, 0, s>>>(L.mem, m, cstart, vel, thq, out, sx,
                                                                 sy, sz, cx, cy, cz);
}

void launch_column_diag(const Launch& L, int m, int ncol, hipStream_t s, const uint32_t* cstart,
                        const float4* thq, const float4* vel, float4* out) {
    k_column_diag<<<nblocks(ncol), TPB, 0, s>>>(L.mem, m, cstart, thq, vel, out);
}

}  // namespace asph
