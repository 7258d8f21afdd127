// asph_handle.h - the device handle behind the C ABI (asph_t) and the step pieces the run
// loops and the adjoint driver share. Host-only; included by asph_host.cpp and
// asph_adjoint_host.cpp.
#pragma once

#include <algorithm>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "asph_host_common.h"

using namespace asph;

struct asph_t {
    std::vector<asph_config> cfgs;
    std::vector<MemberDesc> mem_h;
    std::vector<BlockMap> pmap_h, bmap_h;
    MemberDesc* mem_d = nullptr;
    BlockMap* pmap_d = nullptr;
    BlockMap* bmap_d = nullptr;
    Launch L{};
    int S = 0;              // stencil (common to all members)
    int kb = 0;             // key bits of one member
    int end_bit = 0;        // sort bits (member index above the member's key bits)
    uint32_t nkeys = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = -1;        // the device current when the handle was created
    int64_t N = 0, ncol = 0;
    float4 *pos = nullptr, *vel = nullptr, *thq = nullptr, *acc = nullptr;
    float4 *pos2 = nullptr, *vel2 = nullptr, *thq2 = nullptr, *acc2 = nullptr;
    float *rho = nullptr, *rho2 = nullptr, *prs = nullptr, *dthdt = nullptr;
    float2* rinv = nullptr;   // density factors of the momentum pairs, written by the density kernel
    int64_t *rain = nullptr, *rain2 = nullptr, *dep = nullptr, *precip = nullptr;
    unsigned long long* clipped = nullptr;
    uint32_t *keys0 = nullptr, *kA = nullptr, *vA = nullptr, *kB = nullptr, *vB = nullptr;
    uint32_t *bh = nullptr, *bo = nullptr, *dtot = nullptr, *cstart = nullptr;
    uint32_t *keys_sorted = nullptr, *perm = nullptr;
    uint32_t* maxv = nullptr;        // [n_members][3][64]
    asph::StepCtl* ctl_d = nullptr;  // device step controller (asph_run_device)
    float* disp = nullptr;           // drift since the last sort [cells per axis]
    uint32_t* disp_slots = nullptr;  // per-step drift maxima (64 slots per axis)
    double* clock = nullptr;         // device simulated time (graph replay)
    double* audit_buf = nullptr;
    std::vector<void*> allocs;
    bool sorted = false, initialized = false;
    bool dep_pending = false;        // dep holds deposits not yet merged into the rain grid
    double simt = 0.0, since_sort = 1e9;
    float sort_dt = 0.5f;
    // hipGraph cache: one executable per (buffer assignment, sort flag, dt); tuning changes drop it
    std::map<std::tuple<void*, void*, void*, void*, void*, void*, int, float>, hipGraphExec_t> graphs;
    // profiling
    bool prof = false;
    unsigned long long* stats = nullptr;   // work counters of the neighbour kernels (enabled)
    float2* rec_s12 = nullptr;             // adjoint tape hooks (null outside the adjoint driver)
    float4* rec_av = nullptr;
    unsigned long long* stats_buf = nullptr;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_live[ASPH_K_COUNT];

    int n_members() const { return (int)cfgs.size(); }
    template <class T> int alloc(T** p, size_t count) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return fail("hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        allocs.push_back(q);
        *p = (T*)q;
        return ASPH_OK;
    }
    hipEvent_t get_event() {
        if (!ev_pool.empty()) { hipEvent_t e = ev_pool.back(); ev_pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};


namespace asph_host {

MomArgs mom_args(asph_t* h, float dt, int mode);
void do_sort(asph_t* h, bool launch);
void do_kick_drift(asph_t* h, float dt, int clock_mode, bool launch);
void do_density_eos(asph_t* h, float dt, double simt, int bits, bool use_clock, bool launch);
void do_momentum_kdk(asph_t* h, float dt, bool launch);
void do_rain(asph_t* h, float dt, bool launch);
void issue_step(asph_t* h, float dt, double simt, bool sort, bool graph, bool launch);
int reset_aux(asph_t* h, bool zero_acc);
int read_maxv(asph_t* h, int m0, int m1, float out[3]);

}  // namespace asph_host
