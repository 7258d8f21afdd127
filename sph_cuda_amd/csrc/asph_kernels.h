// asph_kernels.h - host-side launchers of the gfx950 kernels (asph_kernels.hip, asph_sort.hip,
// asph_nbr.hip).
#pragma once

#include "asph_internal.h"

namespace asph {

// Device-resident step controller (asph_run_device): the loop state that asph_run keeps on the
// host, read and advanced on the device inside a captured graph. Kernels given one take dt from
// it and do nothing while it is idle (stopped, paused for a dump, or aborted); the sort stages
// run only on a sort step, the gather then copying without permuting so that the buffer rotation
// captured in the graph stays the same.
struct StepCtl {
    double simt, since_sort;          // start of the current step; time since the last sort
    double sim_end, next_dump, dump_every;
    int64_t steps, sorts, max_steps;
    int64_t n;                        // particles (idle steps copy the rotated buffers)
    float dt, dt_required, sort_dt, clamp_dt;   // clamp_dt: dt adopted after a velocity clamp
    float u[3];
    int32_t do_sort, idle, clamp, trips, abort, pause;
    float dt_cap, dt_floor, cfl, cs, dx, dy, dz;
    int32_t max_trip, strict, every, fixed;
};

// Device tables describing one handle's members and the work maps of its launches.
struct Launch {
    const MemberDesc* mem;   // [n_members]
    const BlockMap* pmap;    // [npblk]: blocks of 256 particles (= cells) of one member each
    const BlockMap* bmap;    // [nbblk]: 8x8x8-cell bricks of one member each
    int npblk;
    int nbblk;
    int n_members;
    // work items of the brick kernels, rebuilt after every sort (launch_brick_items)
    int2* items;             // [nitem_max]: (brick, chunk), heaviest first
    uint32_t* nitems;        // number of valid items
    uint32_t* item_hist;     // [brick_item_hist_words(nbblk)] counting-sort scratch
    uint32_t* item_meta;     // [nbblk] scratch
    int nitem_max;           // grid of the brick kernels (bound for the smallest chunk)
    int chunk;               // listed particles per work item, at most (tuning)
    int order;               // 1: items heaviest first, 0: brick (key) order (tuning)
    // runs of empty keys longer than cell_gap_serial(), filled by whole workgroups
    uint4* gaps;             // [gapcap]: (first key, last key, value)
    uint32_t* ngaps;         // reset by the keys kernel of each sort
    uint32_t gapcap;
    const StepCtl* ctl = nullptr;   // device controller of the launches (asph_run_device), or null
    // drift bound per brick: how far the particles listed in a brick lie beyond their listed cells
    float4* wexc = nullptr;  // [npblk * particle_block_size() / 64]: per-wavefront maxima (kick/drift)
    float4* bexc = nullptr;  // [nkeys >> brick_bits()]: per-brick maxima, indexed by key >> brick_bits()
    float4* bhalo = nullptr; // [nbblk]: per brick of bmap, the maxima over the bricks its halo can reach
};

struct DensArgs {
    const float4* pos;
    const float* rho_old;
    float* rho_new;
    float4* thq;
    float* prs;
    float4* vel;                 // vel.w receives p' for the momentum kernel
    const uint32_t* cstart;
    const uint32_t* keys_sorted;
    const int64_t* rain;
    int64_t* dep;
    uint32_t* maxv;              // [n_members][3][64], cleared here for the momentum kernel
    const float* disp;
    const uint32_t* disp_slots;
    const float4* bhalo = nullptr;  // drift bound per brick of bmap (Launch::bhalo), or null: the member-wide one
    const double* clock;         // simulated time at the start of the step (device), or null:
    double simt;                 // then this value is used
    unsigned long long* stats;   // work counters (ASPH_C_DENS_*), or null
    float2* s12_out;             // adjoint tape: the density sums (S1, S2) per particle, or null
    float2* rinv;                // (max(rho_new, 1e-3), its reciprocal) for the momentum kernel
    const StepCtl* ctl = nullptr;   // device controller (dt, idle), or null
    float dt;
    int bits;                    // bit 0: EOS + microphysics, bit 1: surface flux (if flagged)
};

struct MomArgs {
    const float4* pos;
    const float4* vel;           // (vx, vy, vz, p')
    float4* thq;
    const float* rho;
    const float2* rinv;          // (max(rho, 1e-3), its reciprocal), from the density kernel
    const float* prs;
    const uint32_t* cstart;
    float4* acc;
    float* dthdt;
    float4* vel_out;
    uint32_t* maxv;              // [n_members][3][64]
    const float* disp;
    const uint32_t* disp_slots;
    const float4* bhalo = nullptr;  // drift bound per brick of bmap (Launch::bhalo), or null: the member-wide one
    unsigned long long* stats;   // work counters (ASPH_C_MOM_*), or null
    float4* av_out;              // adjoint tape: the viscous acceleration sums (x mm), or null
    const StepCtl* ctl = nullptr;   // device controller (dt, idle), or null
    float dt;
    int mode;                    // 0: force evaluation, 1: step
};

// ---- adjoint (asph_adjoint.hip) ----------------------------------------------------------------
// Forward values of one step, recorded while a segment is recomputed (device pointers).
struct AdjTape {
    const float4 *x0, *v0, *a0, *t0;   // state at the start of the step
    const float* r0;                   // density at the start of the step
    const int64_t* rain0;              // rain grid at the start of the step
    const float4 *x1, *v1;             // after kick + drift; v1.w = p'
    const float* r1;                   // density of the step
    const float2* s12;                 // density sums (S1, S2)
    const float4* t1;                  // theta, qv, qc after the EOS stage
    const float4 *a1, *av;             // acceleration and viscous sums (x mm) of the step
    const uint32_t* keys;              // sorted keys of the segment (listed cells)
    const uint32_t* cstart;            // cell start table of the segment
    const float* delta;                // drift bound of the step's neighbour stages [cells per axis]
};
// Adjoints of the state in the particle order of a segment.
struct AdjState {
    float4 *x, *v, *a, *t;
    float* rho;
    double* rain;                      // per cell
};
// Adjoints internal to one step's reverse.
struct AdjWork {
    float4 *Ab, *AVb, *x1b, *v1b, *t2b;
    float *pb, *r1b;
    float2* sb;
    double* rb_cur;
};
// mode 1: a KDK step; mode 0: the force evaluation at t = 0 (no motion, no flux, rain merge).
void launch_adj_step(const Launch& L, int S, const DevParams* P, int64_t N, hipStream_t s, const AdjTape& T,
                     const AdjWork& W, const AdjState& E, const AdjState& Sst, double* qc0b, float dt, float Ts,
                     float qs_sfc, float frac, int mode);
// The reverse pair stages (kind 0: momentum, 1: density) on the brick decomposition of the
// segment's sort (asph_nbr.hip).
void launch_adj_pairs(const Launch& L, int S, hipStream_t s, int kind, const AdjTape& T, const AdjWork& W,
                      const AdjState& Sst);
void launch_adj_delta(hipStream_t s, const float* disp, const uint32_t* slots, float* out);
// Up to 8 device-to-device copies in one launch (sizes in 4-byte words; 16-byte aligned buffers).
struct AdjCopies {
    const uint32_t* src[8];
    uint32_t* dst[8];
    uint32_t words[8];
    int count = 0;
    template <class T> void add(T* d, const T* s, size_t n) {
        src[count] = (const uint32_t*)s;
        dst[count] = (uint32_t*)d;
        words[count] = (uint32_t)(n * sizeof(T) / 4);
        ++count;
    }
};
void launch_adj_copies(hipStream_t s, const AdjCopies& C);
void launch_adj_unperm(int64_t N, hipStream_t s, const uint32_t* perm, const AdjState& in, const AdjState& out);

int particle_block_size();
void launch_step_begin(hipStream_t s, float* disp, uint32_t* disp_slots, int reset, double* clock,
                       int clock_mode, double t, float dt, const StepCtl* ctl = nullptr, int sort_gate = 0);
// The controller's steps: before a step (stop conditions, the sort decision) and after it (time,
// CFL rules, velocity clamp flag, abort, sort interval, dump pause; asph_h100.cu:1038-1081).
void launch_ctl_pre(hipStream_t s, StepCtl* ctl);
void launch_ctl_post(hipStream_t s, StepCtl* ctl, const uint32_t* maxv);
void launch_clamp_vel_ctl(const Launch& L, hipStream_t s, float4* vel);
// Kick and drift. With the sorted keys and the cell table of the current sort (keys non-null), it
// also records, per brick, how far the particles listed in it now lie beyond their listed cells
// (L.wexc, then L.bexc).
void launch_kick_drift(const Launch& L, hipStream_t s, float4* pos, float4* vel, const float4* acc,
                       int64_t* dep, uint32_t* disp_slots, float dt, const uint32_t* keys,
                       const uint32_t* cstart);
// The drift bound of each brick's work items from the per-brick bounds (after launch_kick_drift).
void launch_brick_halo(const Launch& L, int S, hipStream_t s);
void launch_density_eos(const Launch& L, int S, hipStream_t s, const DensArgs& A);
void launch_momentum_kdk(const Launch& L, int S, hipStream_t s, const MomArgs& A);
int brick_bits();
size_t brick_item_hist_words(int nbblk);
int brick_item_capacity(int nbblk, int64_t n);
int brick_chunk_min();
int brick_chunk_max();
void launch_brick_items(const Launch& L, hipStream_t s, const uint32_t* cstart);
void launch_rain(const Launch& L, hipStream_t s, const int64_t* rain_in, const int64_t* dep,
                 int64_t* rain_out, int64_t* precip, unsigned long long* clipped, float dt);
void launch_rain_merge(hipStream_t s, int n, int64_t* rain, int64_t* dep);
void launch_keys(const Launch& L, hipStream_t s, const float4* pos, uint32_t* keys);
void launch_gather(hipStream_t s, int N, const uint32_t* perm, const float4* pA, float4* pB,
                   const float4* vA, float4* vB, const float4* tA, float4* tB, const float* rA,
                   float* rB, const float4* aA, float4* aB, const StepCtl* ctl = nullptr);
uint32_t cell_gap_capacity(uint32_t nkeys);
void launch_cell_start(const Launch& L, hipStream_t s, int N, uint32_t nkeys, const uint32_t* keys,
                       uint32_t* cstart);
void launch_clamp_vel(const Launch& L, hipStream_t s, float4* vel, float vmax, float cs_frac);
void launch_audit(const Launch& L, hipStream_t s, const float4* pos, const float4* vel,
                  const float4* thq, const int64_t* rain, const int64_t* dep, const int64_t* precip,
                  double* tot);
void launch_coarse_dump(const Launch& L, int m, hipStream_t s, const uint32_t* cstart,
                        const float4* vel, const float4* thq, float* out, int sx, int sy, int sz,
                        int cx, int cy, int cz);
void launch_column_diag(const Launch& L, int m, int ncol, hipStream_t s, const uint32_t* cstart,
                        const float4* thq, const float4* vel, float4* out);

int radix_tile();
// Stable LSD radix sort of the keys in keys0 (bits [0, end_bit)) with their indices.
// (kA,vA) and (kB,vB) are ping-pong buffers; bh, bo hold 256*nblk counts, dtot 256 totals.
// On return *keys_sorted and *perm point into kA/vA or kB/vB.
void launch_radix_sort(hipStream_t s, int N, int end_bit, const uint32_t* keys0, uint32_t* kA,
                       uint32_t* vA, uint32_t* kB, uint32_t* vB, uint32_t* bh, uint32_t* bo,
                       uint32_t* dtot, uint32_t** keys_sorted, uint32_t** perm, const StepCtl* ctl = nullptr);

}  // namespace asph
