// asph_sort.hip - Morton keys and a stable LSD radix sort of (key, index) pairs for gfx950.
//
// Replaces kernel_iota + cub::DeviceRadixSort::SortPairs (asph_h100.cu:241-244, 869, 903, 966):
// the result is the stable argsort of the keys over bits [0, end_bit), i.e. ties keep the
// original index order. The keys come from k_keys (asph_kernels.hip).
//
// Each pass sorts one 8-bit digit (the last pass may be narrower) in three steps:
//   1. k_radix_hist: each workgroup counts the digits of one tile of RS_TILE keys in LDS
//      (wave-aggregated: lanes holding the same digit are found with 8 ballots and one lane
//      adds the group's size), and stores its 256 counts digit-major in bh[d * nblk + b].
//   2. k_radix_scan_digits / k_radix_scan_totals: exclusive scan of bh in (digit, tile) order.
//   3. k_radix_scatter: each workgroup ranks its tile stably in LDS -- rounds of 256 keys in
//      index order; within a wave the rank is the popcount of the equal-digit lanes below,
//      across waves and rounds it is carried in LDS counters -- places the keys in LDS sorted
//      by digit, then writes each digit's run to its global offset, so consecutive lanes write
//      consecutive addresses.
#include "asph_kernels.h"

namespace asph {

namespace {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;

// Mask of the lanes of this wave whose digit equals ours (bits = digit width).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, int bits, bool valid) {
    uint64_t m = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
        uint64_t bal = __ballot((d >> b) & 1u);
        m &= ((d >> b) & 1u) ? bal : ~bal;
    }
    return m;
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const uint32_t* __restrict__ kin, int N,
                                                           int shift, int bits,
                                                           uint32_t* __restrict__ bh, int nblk, const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    __shared__ uint32_t h[256];
    const int t = threadIdx.x;
    const int lane = t & 63;
    h[t] = 0;
    __syncthreads();
    const uint32_t mask = (1u << bits) - 1u;
    const int base = blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        int idx = base + r * RS_THREADS + t;
        bool valid = idx < N;
        uint32_t k = valid ? kin[idx] : 0u;
        uint32_t d = (k >> shift) & mask;
        uint64_t m = match_digit(d, bits, valid);
        if (valid && lane == __ffsll((unsigned long long)m) - 1) atomicAdd(&h[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    bh[(size_t)t * nblk + blockIdx.x] = h[t];
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& excl, uint32_t* s_w) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t woff = 0, total = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; ++k) {
        uint32_t sw = s_w[k];
        if (k < w) woff += sw;
        total += sw;
    }
    __syncthreads();
    excl = woff + x - v;
    return total;
}

// One workgroup per digit: exclusive scan of bh[d*nblk .. d*nblk+nblk) into bo, total to dtot.
__global__ __launch_bounds__(RS_THREADS) void k_radix_scan_digits(const uint32_t* __restrict__ bh,
                                                                  uint32_t* __restrict__ bo,
                                                                  uint32_t* __restrict__ dtot,
                                                                  int nblk, const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    __shared__ uint32_t s_w[RS_WAVES];
    const size_t row = (size_t)blockIdx.x * nblk;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < nblk; c0 += RS_THREADS * 4) {
        // 4 consecutive entries per thread
        uint32_t v[4], sum = 0;
        int b0 = c0 + threadIdx.x * 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = (b0 + k < nblk) ? bh[row + b0 + k] : 0u;
            sum += v[k];
        }
        uint32_t excl;
        uint32_t total = block_exclusive_scan(sum, excl, s_w);
        uint32_t run = carry + excl;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (b0 + k < nblk) bo[row + b0 + k] = run;
            run += v[k];
        }
        carry += total;
    }
    if (threadIdx.x == 0) dtot[blockIdx.x] = carry;
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scan_totals(uint32_t* __restrict__ dtot, const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    __shared__ uint32_t s_w[RS_WAVES];
    uint32_t v = dtot[threadIdx.x], excl;
    block_exclusive_scan(v, excl, s_w);
    dtot[threadIdx.x] = excl;
}

__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int N, int shift, int bits, const uint32_t* __restrict__ bh,
    const uint32_t* __restrict__ bo, const uint32_t* __restrict__ dbase, int nblk, const StepCtl* __restrict__ ctl) {
    if (ctl && !ctl->do_sort) return;
    __shared__ uint32_t s_key[RS_TILE];
    __shared__ uint32_t s_val[RS_TILE];
    __shared__ uint32_t s_cnt[256];
    __shared__ uint32_t s_wcnt[RS_WAVES][256];
    __shared__ uint32_t s_tstart[256];
    __shared__ uint32_t s_gbase[256];
    __shared__ uint32_t s_w[RS_WAVES];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int b = blockIdx.x;
    const uint32_t mask = (1u << bits) - 1u;
    const int base = b * RS_TILE;
    const int count = min(RS_TILE, N - base);

    uint32_t c = bh[(size_t)t * nblk + b];
    uint32_t excl;
    block_exclusive_scan(c, excl, s_w);
    s_tstart[t] = excl;
    s_gbase[t] = dbase[t] + bo[(size_t)t * nblk + b];
    s_cnt[t] = 0;
#pragma unroll
    for (int k = 0; k < RS_WAVES; ++k) s_wcnt[k][t] = 0;
    __syncthreads();

    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for (int r = 0; r < RS_ITEMS; ++r) {
        int li = r * RS_THREADS + t;
        bool valid = li < count;
        uint32_t k = 0, v = 0;
        if (valid) {
            k = kin[base + li];
            v = vin ? vin[base + li] : (uint32_t)(base + li);
        }
        uint32_t d = (k >> shift) & mask;
        uint64_t m = match_digit(d, bits, valid);
        uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && lane == __ffsll((unsigned long long)m) - 1) s_wcnt[w][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t off = s_cnt[d] + rank;
            for (int k2 = 0; k2 < w; ++k2) off += s_wcnt[k2][d];
            uint32_t p = s_tstart[d] + off;
            s_key[p] = k;
            s_val[p] = v;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int k2 = 0; k2 < RS_WAVES; ++k2) {
            add += s_wcnt[k2][t];
            s_wcnt[k2][t] = 0;
        }
        s_cnt[t] += add;
        __syncthreads();
    }
    for (int p = t; p < count; p += RS_THREADS) {
        uint32_t k = s_key[p];
        uint32_t d = (k >> shift) & mask;
        uint32_t g = s_gbase[d] + (uint32_t)p - s_tstart[d];
        kout[g] = k;
        vout[g] = s_val[p];
    }
}

}  // namespace

int radix_tile() { return RS_TILE; }

void launch_radix_sort(hipStream_t s, int N, int end_bit, const uint32_t* keys0, uint32_t* kA,
                       uint32_t* vA, uint32_t* kB, uint32_t* vB, uint32_t* bh, uint32_t* bo,
                       uint32_t* dtot, uint32_t** keys_sorted, uint32_t** perm, const StepCtl* ctl) {
    const int nblk = (N + RS_TILE - 1) / RS_TILE;
    const int npass = (end_bit + 7) / 8;
    const uint32_t* kin = keys0;
    const uint32_t* vin = nullptr;
    uint32_t* ko = kA;
    uint32_t* vo = vA;
    for (int p = 0; p < npass; ++p) {
        int shift = 8 * p;
        int bits = min(8, end_bit - shift);
        k_radix_hist<<<nblk, RS_THREADS, 0, s>>>(kin, N, shift, bits, bh, nblk, ctl);
        k_radix_scan_digits<<<256, RS_THREADS, 0, s>>>(bh, bo, dtot, nblk, ctl);
        k_radix_scan_totals<<<1, RS_THREADS, 0, s>>>(dtot, ctl);
        k_radix_scatter<<<nblk, RS_THREADS, 0, s>>>(kin, vin, ko, vo, N, shift, bits, bh, bo, dtot,
                                                    nblk, ctl);
        kin = ko;
        vin = vo;
        if (ko == kA) { ko = kB; vo = vB; } else { ko = kA; vo = vA; }
    }
    *keys_sorted = const_cast<uint32_t*>(kin);
    *perm = const_cast<uint32_t*>(vin);
}

}  // namespace asph
