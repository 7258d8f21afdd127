// asph_host.cpp - the device side of the C ABI (include/asph.h): handles (a single domain or an
// ensemble of domains), state transfer, step orchestration, the reference's closed-loop CFL run
// loop, hipGraph replay of whole steps, diagnostics and the dump writer.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "asph_host_common.h"

using namespace asph;

#include "asph_handle.h"

namespace {

bool g_capturing = false;   // no profiling events inside a graph capture

// Brackets a launch with profiling events when enabled.
struct ProfScope {
    asph_t* h;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(asph_t* h_, int k_) : h(h_), k(k_) {
        if (h->prof && !g_capturing) {
            a = h->get_event();
            b = h->get_event();
            if (a) (void)hipEventRecord(a, h->stream);
        }
    }
    ~ProfScope() {
        if (a && b) {
            (void)hipEventRecord(b, h->stream);
            h->ev_live[k].push_back({a, b});
        }
    }
};

int bits_for(int n) { int b = 0; while ((1 << b) < n) ++b; return b; }

// Builds the member table and the launch maps of a handle.
int build_tables(asph_t* h) {
    const int nm = h->n_members();
    h->kb = 0;
    h->S = stencil_of(h->cfgs[0]);
    for (auto& c : h->cfgs) {
        DevParams P;
        make_params(c, P);
        h->kb = std::max(h->kb, P.end_bit);
        if (stencil_of(c) != h->S) return fail("all members of an ensemble need the same stencil");
    }
    h->end_bit = h->kb + bits_for(nm);
    if (h->end_bit > 31) return fail("ensemble key space needs %d bits (> 31)", h->end_bit);
    h->nkeys = 1u << h->end_bit;
    const int TPB = particle_block_size();
    const int BSZ = 1 << brick_bits();
    const int BE = 8;   // brick edge
    int64_t p0 = 0, col0 = 0;
    h->mem_h.clear(); h->pmap_h.clear(); h->bmap_h.clear();
    for (int m = 0; m < nm; ++m) {
        MemberDesc d;
        std::memset(&d, 0, sizeof(d));
        make_params(h->cfgs[m], d.P);
        d.p0 = (int32_t)p0;
        d.col0 = (int32_t)col0;
        d.k0 = (uint32_t)m << h->kb;
        d.first_block = (int32_t)h->pmap_h.size();
        for (int b = 0; b < d.P.N; b += TPB) h->pmap_h.push_back({m, b, 0, 0});
        // bricks covering the member's cell box, in Morton order
        std::vector<uint32_t> keys;
        for (int bz = 0; bz < d.P.nz; bz += BE)
            for (int by = 0; by < d.P.ny; by += BE)
                for (int bx = 0; bx < d.P.nx; bx += BE) keys.push_back(morton3(bx, by, bz));
        std::sort(keys.begin(), keys.end());
        for (size_t k = 0; k < keys.size(); ++k) h->bmap_h.push_back({m, (int32_t)keys[k], (int32_t)k, 0});
        if ((int64_t)keys.size() * BSZ < d.P.N) return fail("internal: bricks do not cover member %d", m);
        h->mem_h.push_back(d);
        p0 += d.P.N;
        col0 += d.P.ncol;
        if (p0 > 0x7fffffffLL) return fail("more than 2^31 particles");
    }
    h->N = p0;
    h->ncol = col0;
    return ASPH_OK;
}

int create_impl(const asph_config* cfgs, int n, void* stream, asph_t** out) {
    if (!out) return fail("null output");
    *out = nullptr;
    if (!cfgs || n < 1) return fail("need at least one member");
    for (int m = 0; m < n; ++m)
        if (asph_config_check(&cfgs[m]) != ASPH_OK) return fail("member %d: %s", m, g_err.c_str());
    asph_t* h = new asph_t();
    h->cfgs.assign(cfgs, cfgs + n);
    if (build_tables(h) != ASPH_OK) { std::string e = g_err; delete h; g_err = e; return ASPH_ERR; }
    if (stream) {
        h->stream = (hipStream_t)stream;
    } else {
        if (hipStreamCreate(&h->stream) != hipSuccess) { delete h; return fail("hipStreamCreate failed"); }
        h->own_stream = true;
    }
    if (hipGetDevice(&h->device) != hipSuccess) h->device = -1;
    const size_t N = (size_t)h->N;
    const int nblk = (int)((N + radix_tile() - 1) / radix_tile());
    int rc = ASPH_OK;
    rc |= h->alloc(&h->pos, N); rc |= h->alloc(&h->vel, N); rc |= h->alloc(&h->thq, N);
    rc |= h->alloc(&h->acc, N); rc |= h->alloc(&h->pos2, N); rc |= h->alloc(&h->vel2, N);
    rc |= h->alloc(&h->thq2, N); rc |= h->alloc(&h->acc2, N);
    rc |= h->alloc(&h->rho, N); rc |= h->alloc(&h->rho2, N); rc |= h->alloc(&h->prs, N);
    rc |= h->alloc(&h->rinv, N);
    rc |= h->alloc(&h->dthdt, N);
    rc |= h->alloc(&h->rain, N); rc |= h->alloc(&h->rain2, N);
    rc |= h->alloc(&h->dep, N); rc |= h->alloc(&h->precip, (size_t)h->ncol);
    rc |= h->alloc(&h->clipped, 1);
    rc |= h->alloc(&h->keys0, N); rc |= h->alloc(&h->kA, N); rc |= h->alloc(&h->vA, N);
    rc |= h->alloc(&h->kB, N); rc |= h->alloc(&h->vB, N);
    rc |= h->alloc(&h->bh, (size_t)256 * nblk); rc |= h->alloc(&h->bo, (size_t)256 * nblk);
    rc |= h->alloc(&h->dtot, 256);
    rc |= h->alloc(&h->cstart, (size_t)h->nkeys + 1);
    rc |= h->alloc(&h->maxv, (size_t)192 * n);
    rc |= h->alloc(&h->disp, 4);
    rc |= h->alloc(&h->disp_slots, 3 * 64);
    float4 *wexc = nullptr, *bexc = nullptr, *bhalo = nullptr;
    const size_t nbexc = (size_t)(h->nkeys >> brick_bits());
    rc |= h->alloc(&wexc, h->pmap_h.size() * (size_t)(particle_block_size() / 64));
    rc |= h->alloc(&bexc, nbexc);
    rc |= h->alloc(&bhalo, h->bmap_h.size());
    rc |= h->alloc(&h->clock, 1);
    rc |= h->alloc(&h->audit_buf, 8);
    rc |= h->alloc(&h->mem_d, h->mem_h.size());
    rc |= h->alloc(&h->pmap_d, h->pmap_h.size());
    rc |= h->alloc(&h->bmap_d, h->bmap_h.size());
    const int nitem_max = brick_item_capacity((int)h->bmap_h.size(), h->N);
    int2* items = nullptr;
    uint32_t *nitems = nullptr, *item_hist = nullptr, *item_meta = nullptr;
    rc |= h->alloc(&items, (size_t)nitem_max);
    rc |= h->alloc(&nitems, 1);
    rc |= h->alloc(&item_hist, brick_item_hist_words((int)h->bmap_h.size()));
    rc |= h->alloc(&item_meta, h->bmap_h.size());
    const uint32_t gapcap = cell_gap_capacity(h->nkeys);
    uint4* gaps = nullptr;
    uint32_t* ngaps = nullptr;
    rc |= h->alloc(&gaps, gapcap);
    rc |= h->alloc(&ngaps, 1);
    if (rc != ASPH_OK) { std::string e = g_err; asph_destroy(h); g_err = e; return ASPH_ERR; }
    hipError_t e = hipMemcpy(h->mem_d, h->mem_h.data(), h->mem_h.size() * sizeof(MemberDesc), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->pmap_d, h->pmap_h.data(), h->pmap_h.size() * sizeof(BlockMap), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->bmap_d, h->bmap_h.data(), h->bmap_h.size() * sizeof(BlockMap), hipMemcpyHostToDevice);
    // no items before a sort
    if (e == hipSuccess) e = hipMemset(nitems, 0, sizeof(uint32_t));
    // per-brick drift bounds: zero holds for a state that has not moved since its sort
    if (e == hipSuccess) e = hipMemset(bexc, 0, nbexc * sizeof(float4));
    if (e == hipSuccess) e = hipMemset(bhalo, 0, h->bmap_h.size() * sizeof(float4));
    if (e != hipSuccess) { asph_destroy(h); return fail("table upload: %s", hipGetErrorString(e)); }
    h->L.mem = h->mem_d; h->L.pmap = h->pmap_d; h->L.bmap = h->bmap_d;
    h->L.npblk = (int)h->pmap_h.size(); h->L.nbblk = (int)h->bmap_h.size(); h->L.n_members = n;
    h->L.items = items; h->L.nitems = nitems; h->L.item_hist = item_hist; h->L.item_meta = item_meta;
    h->L.nitem_max = nitem_max;
    h->L.wexc = wexc; h->L.bexc = bexc; h->L.bhalo = bhalo;
    h->L.chunk = 768;
    h->L.order = 1;
    h->L.gaps = gaps; h->L.ngaps = ngaps; h->L.gapcap = gapcap;
    *out = h;
    return ASPH_OK;
}

}  // namespace

// Step pieces shared with the adjoint driver (asph_adjoint_host.cpp).
namespace asph_host {

MomArgs mom_args(asph_t* h, float dt, int mode) {
    MomArgs A;
    A.pos = h->pos; A.vel = h->vel; A.thq = h->thq; A.rho = h->rho; A.rinv = h->rinv; A.prs = h->prs;
    A.cstart = h->cstart; A.acc = h->acc; A.dthdt = h->dthdt; A.vel_out = h->vel2;
    A.maxv = h->maxv; A.disp = h->disp; A.disp_slots = h->disp_slots;
    A.bhalo = h->L.bhalo;
    A.stats = h->stats;
    A.av_out = h->rec_av;
    A.ctl = h->L.ctl;
    A.dt = dt; A.mode = mode;
    return A;
}

// --- step pieces (launch = false only replays the buffer swaps of a captured graph) ---------
void do_sort(asph_t* h, bool launch) {
    if (launch) {
        ProfScope ps(h, ASPH_K_SORT);
        launch_keys(h->L, h->stream, h->pos, h->keys0);
        launch_radix_sort(h->stream, (int)h->N, h->end_bit, h->keys0, h->kA, h->vA, h->kB, h->vB,
                          h->bh, h->bo, h->dtot, &h->keys_sorted, &h->perm, h->L.ctl);
        const bool ga = (h->cfgs[0].flags & ASPH_F_GATHER_ACC) != 0;
        launch_gather(h->stream, (int)h->N, h->perm, h->pos, h->pos2, h->vel, h->vel2, h->thq,
                      h->thq2, h->rho, h->rho2, ga ? h->acc : nullptr, ga ? h->acc2 : nullptr, h->L.ctl);
        launch_cell_start(h->L, h->stream, (int)h->N, h->nkeys, h->keys_sorted, h->cstart);
        launch_brick_items(h->L, h->stream, h->cstart);
        launch_step_begin(h->stream, h->disp, h->disp_slots, 1, h->clock, 2, 0.0, 0.f, h->L.ctl, 1);
    } else {
        // the sorted keys and permutation land in A or B depending on the pass count
        const int npass = (h->end_bit + 7) / 8;
        h->keys_sorted = (npass & 1) ? h->kA : h->kB;
        h->perm = (npass & 1) ? h->vA : h->vB;
    }
    std::swap(h->pos, h->pos2); std::swap(h->vel, h->vel2); std::swap(h->thq, h->thq2);
    std::swap(h->rho, h->rho2);
    if (h->cfgs[0].flags & ASPH_F_GATHER_ACC) std::swap(h->acc, h->acc2);
    h->sorted = true;
}

void do_kick_drift(asph_t* h, float dt, int clock_mode, bool launch) {
    if (launch) {
        ProfScope ps(h, ASPH_K_KICK_DRIFT);
        launch_step_begin(h->stream, h->disp, h->disp_slots, 0, h->clock, clock_mode, 0.0, dt, h->L.ctl, 0);
        // the per-brick drift bound needs the listing of a sort (else the next sort resets it)
        launch_kick_drift(h->L, h->stream, h->pos, h->vel, h->acc, h->dep, h->disp_slots, dt,
                          h->sorted ? h->keys_sorted : nullptr, h->sorted ? h->cstart : nullptr);
        if (h->sorted) launch_brick_halo(h->L, h->S, h->stream);
    }
    h->dep_pending = false;
}

void do_density_eos(asph_t* h, float dt, double simt, int bits, bool use_clock, bool launch) {
    if (launch) {
        DensArgs A;
        A.pos = h->pos; A.rho_old = h->rho; A.rho_new = h->rho2; A.thq = h->thq; A.prs = h->prs;
        A.vel = h->vel; A.cstart = h->cstart; A.keys_sorted = h->keys_sorted; A.rain = h->rain;
        A.dep = h->dep; A.maxv = h->maxv; A.disp = h->disp; A.disp_slots = h->disp_slots;
        A.bhalo = h->L.bhalo;
        A.clock = use_clock ? h->clock : nullptr; A.simt = simt;
        if (h->L.ctl) A.clock = &h->L.ctl->simt;   // device controller: the step's start time
        A.ctl = h->L.ctl;
        A.stats = h->stats;
        A.s12_out = h->rec_s12;
        A.rinv = h->rinv;
        A.dt = dt; A.bits = bits;
        ProfScope ps(h, ASPH_K_DENSITY_EOS);
        launch_density_eos(h->L, h->S, h->stream, A);
    }
    std::swap(h->rho, h->rho2);
    if (bits & 1) h->dep_pending = true;
}

void do_momentum_kdk(asph_t* h, float dt, bool launch) {
    if (launch) {
        ProfScope ps(h, ASPH_K_MOMENTUM_KDK);
        launch_momentum_kdk(h->L, h->S, h->stream, mom_args(h, dt, 1));
    }
    std::swap(h->vel, h->vel2);
}

void do_rain(asph_t* h, float dt, bool launch) {
    if (launch) {
        ProfScope ps(h, ASPH_K_RAIN);
        launch_rain(h->L, h->stream, h->rain, h->dep, h->rain2, h->precip, h->clipped, dt);
    }
    std::swap(h->rain, h->rain2);
    h->dep_pending = false;
}

// One KDK step in the reference order. clock_mode 1 advances the device clock (graph replay).
void issue_step(asph_t* h, float dt, double simt, bool sort, bool graph, bool launch) {
    if (sort) do_sort(h, launch);
    do_kick_drift(h, dt, graph ? 1 : 2, launch);
    do_density_eos(h, dt, simt, 1 | 2, graph, launch);
    do_momentum_kdk(h, dt, launch);
    do_rain(h, dt, launch);
}

int reset_aux(asph_t* h, bool zero_acc) {
    const size_t N = (size_t)h->N;
    if (zero_acc) HIPCHK(hipMemsetAsync(h->acc, 0, N * sizeof(float4), h->stream));
    HIPCHK(hipMemsetAsync(h->prs, 0, N * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->dthdt, 0, N * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->rain, 0, N * sizeof(int64_t), h->stream));
    HIPCHK(hipMemsetAsync(h->dep, 0, N * sizeof(int64_t), h->stream));
    HIPCHK(hipMemsetAsync(h->precip, 0, (size_t)h->ncol * sizeof(int64_t), h->stream));
    HIPCHK(hipMemsetAsync(h->clipped, 0, sizeof(unsigned long long), h->stream));
    HIPCHK(hipMemsetAsync(h->maxv, 0, (size_t)192 * h->n_members() * sizeof(uint32_t), h->stream));
    HIPCHK(hipMemsetAsync(h->disp, 0, 4 * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->disp_slots, 0, 3 * 64 * sizeof(uint32_t), h->stream));
    HIPCHK(hipMemsetAsync(h->L.bexc, 0, (size_t)(h->nkeys >> brick_bits()) * sizeof(float4), h->stream));
    HIPCHK(hipMemsetAsync(h->L.bhalo, 0, h->bmap_h.size() * sizeof(float4), h->stream));
    HIPCHK(hipMemsetAsync(h->clock, 0, sizeof(double), h->stream));
    h->sorted = false;
    h->initialized = false;
    h->dep_pending = false;
    h->simt = 0.0; h->since_sort = 1e9; h->sort_dt = 0.5f;
    return ASPH_OK;
}

int read_maxv(asph_t* h, int m0, int m1, float out[3]) {
    std::vector<uint32_t> v((size_t)192 * (m1 - m0));
    HIPCHK(hipMemcpyAsync(v.data(), h->maxv + (size_t)192 * m0, v.size() * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int a = 0; a < 3; ++a) {
        uint32_t mx = 0;
        for (int m = 0; m < m1 - m0; ++m)
            for (int s = 0; s < 64; ++s) mx = std::max(mx, v[(size_t)192 * m + 64 * a + s]);
        std::memcpy(&out[a], &mx, 4);
    }
    return ASPH_OK;
}

// The sort interval of the reference's gating (asph_h100.cu:1075-1077) from the max speeds.
float sort_interval(const asph_config& c, const float u[3], float dt) {
    float dz_move = 0.3f * c.dz / std::fmax(u[2], 0.5f);
    float dh_move = 0.3f * c.dx / std::fmax(std::fmax(u[0], u[1]), 0.5f);
    return std::fmin(std::fmax(std::fmin(dz_move, dh_move), dt), 15.0f);
}

}  // namespace asph_host

using namespace asph_host;

extern "C" {

size_t asph_sizeof_config(void) { return sizeof(asph_config); }
size_t asph_sizeof_ic_spec(void) { return sizeof(asph_ic_spec); }
size_t asph_sizeof_run_args(void) { return sizeof(asph_run_args); }
size_t asph_sizeof_run_stats(void) { return sizeof(asph_run_stats); }

int asph_create(const asph_config* cfg, void* stream, asph_t** out) {
    return create_impl(cfg, 1, stream, out);
}

int asph_ensemble_create(const asph_config* cfgs, int32_t n_members, void* stream, asph_t** out) {
    return create_impl(cfgs, n_members, stream, out);
}

int32_t asph_member_count(const asph_t* h) { return h ? h->n_members() : 0; }

int asph_member_range(const asph_t* h, int32_t m, int64_t* first, int64_t* count) {
    if (!h || m < 0 || m >= h->n_members()) return fail("bad member index");
    if (first) *first = h->mem_h[m].p0;
    if (count) *count = h->mem_h[m].P.N;
    return ASPH_OK;
}

int asph_destroy(asph_t* h) {
    if (!h) return ASPH_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    for (void* p : h->allocs) (void)hipFree(p);
    for (auto& e : h->ev_pool) (void)hipEventDestroy(e);
    for (auto& v : h->ev_live) for (auto& pr : v) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return ASPH_OK;
}

int asph_get_config(const asph_t* h, asph_config* cfg) {
    if (!h || !cfg) return fail("null argument");
    *cfg = h->cfgs[0];
    return ASPH_OK;
}

int64_t asph_num_particles(const asph_t* h) { return h ? h->N : 0; }
int asph_sort_end_bit(const asph_t* h) { return h ? h->end_bit : 0; }

int asph_set_state(asph_t* h, const float* pos4, const float* vel4, const float* thq4, const float* rho) {
    if (!h || !pos4 || !vel4 || !thq4 || !rho) return fail("null argument");
    const size_t N = (size_t)h->N;
    HIPCHK(hipMemcpyAsync(h->pos, pos4, N * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->vel, vel4, N * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->thq, thq4, N * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->rho, rho, N * sizeof(float), hipMemcpyHostToDevice, h->stream));
    if (reset_aux(h, true) != ASPH_OK) return ASPH_ERR;
    HIPCHK(hipStreamSynchronize(h->stream));
    return ASPH_OK;
}

int asph_set_state_device(asph_t* h, const float* pos4, const float* vel4, const float* thq4,
                          const float* rho, const float* acc4) {
    if (!h || !pos4 || !vel4 || !thq4 || !rho) return fail("null argument");
    const size_t N = (size_t)h->N;
    HIPCHK(hipMemcpyAsync(h->pos, pos4, N * sizeof(float4), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->vel, vel4, N * sizeof(float4), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->thq, thq4, N * sizeof(float4), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->rho, rho, N * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
    if (acc4) HIPCHK(hipMemcpyAsync(h->acc, acc4, N * sizeof(float4), hipMemcpyDeviceToDevice, h->stream));
    return reset_aux(h, acc4 == nullptr);
}

int asph_get_state(asph_t* h, float* pos4, float* vel4, float* thq4, float* rho, float* acc4,
                   float* prs, float* dthdt) {
    if (!h) return fail("null handle");
    const size_t N = (size_t)h->N;
    if (pos4) HIPCHK(hipMemcpyAsync(pos4, h->pos, N * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    if (vel4) HIPCHK(hipMemcpyAsync(vel4, h->vel, N * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    if (thq4) HIPCHK(hipMemcpyAsync(thq4, h->thq, N * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    if (rho) HIPCHK(hipMemcpyAsync(rho, h->rho, N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    if (acc4) HIPCHK(hipMemcpyAsync(acc4, h->acc, N * sizeof(float4), hipMemcpyDeviceToHost, h->stream));
    if (prs) HIPCHK(hipMemcpyAsync(prs, h->prs, N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    if (dthdt) HIPCHK(hipMemcpyAsync(dthdt, h->dthdt, N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // vel.w carries p' internally; the reference layout has w = 0
    if (vel4) for (size_t i = 0; i < N; ++i) vel4[4 * i + 3] = 0.f;
    return ASPH_OK;
}

int asph_state_device_ptrs(asph_t* h, float** pos4, float** vel4, float** thq4, float** rho, float** acc4) {
    if (!h) return fail("null handle");
    if (pos4) *pos4 = (float*)h->pos;
    if (vel4) *vel4 = (float*)h->vel;
    if (thq4) *thq4 = (float*)h->thq;
    if (rho) *rho = h->rho;
    if (acc4) *acc4 = (float*)h->acc;
    return ASPH_OK;
}

int asph_set_acc(asph_t* h, const float* acc4) {
    if (!h || !acc4) return fail("null argument");
    HIPCHK(hipMemcpyAsync(h->acc, acc4, (size_t)h->N * sizeof(float4), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return ASPH_OK;
}

int asph_get_rain_fx(asph_t* h, int64_t* rain, int64_t* precip, int64_t* clipped) {
    if (!h) return fail("null handle");
    std::vector<int64_t> r, d;
    if (rain) {
        r.resize(h->N); d.resize(h->N);
        HIPCHK(hipMemcpyAsync(r.data(), h->rain, h->N * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(d.data(), h->dep, h->N * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    }
    if (precip) HIPCHK(hipMemcpyAsync(precip, h->precip, h->ncol * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    unsigned long long cl = 0;
    if (clipped) HIPCHK(hipMemcpyAsync(&cl, h->clipped, sizeof(cl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // pending deposits (dep) belong to the grid until the next step consumes them
    if (rain) for (int64_t c = 0; c < h->N; ++c) rain[c] = r[c] + (h->dep_pending ? d[c] : 0);
    if (clipped) *clipped = (int64_t)cl;
    return ASPH_OK;
}

int asph_get_rain(asph_t* h, float* rain_kg, float* precip_kgm2) {
    if (!h) return fail("null handle");
    std::vector<int64_t> r(rain_kg ? h->N : 0), p(precip_kgm2 ? h->ncol : 0);
    if (asph_get_rain_fx(h, rain_kg ? r.data() : nullptr, precip_kgm2 ? p.data() : nullptr, nullptr) != ASPH_OK)
        return ASPH_ERR;
    for (int m = 0; m < h->n_members(); ++m) {
        const MemberDesc& d = h->mem_h[m];
        const double mass = (double)d.P.pmass * FX_INV;
        if (rain_kg)
            for (int64_t c = 0; c < d.P.ncell; ++c) rain_kg[d.p0 + c] = (float)((double)r[d.p0 + c] * mass);
        if (precip_kgm2) {
            const double a = mass / ((double)d.P.dx * (double)d.P.dy);
            for (int64_t c = 0; c < d.P.ncol; ++c) precip_kgm2[d.col0 + c] = (float)((double)p[d.col0 + c] * a);
        }
    }
    return ASPH_OK;
}

int asph_set_rain(asph_t* h, const float* rain_kg) {
    if (!h || !rain_kg) return fail("null argument");
    std::vector<int64_t> r(h->N);
    for (int m = 0; m < h->n_members(); ++m) {
        const MemberDesc& d = h->mem_h[m];
        const double inv = FX_SCALE / (double)d.P.pmass;
        for (int64_t c = 0; c < d.P.ncell; ++c)
            r[d.p0 + c] = (int64_t)std::llrint((double)rain_kg[d.p0 + c] * inv);
    }
    HIPCHK(hipMemcpyAsync(h->rain, r.data(), h->N * sizeof(int64_t), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemsetAsync(h->dep, 0, h->N * sizeof(int64_t), h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->dep_pending = false;
    return ASPH_OK;
}

int asph_sort(asph_t* h) {
    if (!h) return fail("null handle");
    do_sort(h, true);
    LAUNCHCHK();
    return ASPH_OK;
}

int asph_get_sort(asph_t* h, uint32_t* keys, uint32_t* keys_sorted, uint32_t* perm,
                  int32_t* cell_range, int32_t* surf_list, int32_t* surf_count) {
    if (!h) return fail("null handle");
    if (!h->sorted) return fail("no sort has run on this state");
    const size_t N = (size_t)h->N;
    std::vector<uint32_t> k0(N), ks(N), pm(N), cs;
    HIPCHK(hipMemcpyAsync(k0.data(), h->keys0, N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(ks.data(), h->keys_sorted, N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(pm.data(), h->perm, N * 4, hipMemcpyDeviceToHost, h->stream));
    if (cell_range) {
        cs.resize((size_t)h->nkeys + 1);
        HIPCHK(hipMemcpyAsync(cs.data(), h->cstart, cs.size() * 4, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    // per member, in the reference layout: member-local keys, indices and row-major cell ranges
    const uint32_t kmask = (1u << h->kb) - 1u;
    int32_t nsurf = 0;
    for (int m = 0; m < h->n_members(); ++m) {
        const MemberDesc& d = h->mem_h[m];
        const DevParams& P = d.P;
        const int64_t a = d.p0, n = P.N;
        for (int64_t i = 0; i < n; ++i) {
            if (keys) keys[a + i] = k0[a + i] & kmask;
            if (keys_sorted) keys_sorted[a + i] = ks[a + i] & kmask;
            if (perm) perm[a + i] = pm[a + i] - (uint32_t)a;
        }
        if (cell_range) {
            for (int iz = 0; iz < P.nz; ++iz)
                for (int iy = 0; iy < P.ny; ++iy)
                    for (int ix = 0; ix < P.nx; ++ix) {
                        int64_t cid = ((int64_t)iz * P.ny + iy) * P.nx + ix;
                        uint32_t k = d.k0 | morton3((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
                        uint32_t s = cs[k], e = cs[k + 1];
                        cell_range[2 * (a + cid)] = s < e ? (int32_t)(s - a) : 0;
                        cell_range[2 * (a + cid) + 1] = s < e ? (int32_t)(e - a) : 0;
                    }
        }
        int64_t cap = 2 * (int64_t)P.ncol, cnt = 0;
        for (int64_t i = 0; i < n && cnt < cap; ++i)
            if ((ks[a + i] & 0x24924924u & kmask) == 0u) {
                if (surf_list) surf_list[nsurf] = (int32_t)i;
                ++nsurf;
                ++cnt;
            }
    }
    if (surf_count) *surf_count = nsurf;
    return ASPH_OK;
}

int asph_force(asph_t* h, float dt) {
    if (!h) return fail("null handle");
    if (!h->sorted) return fail("asph_sort must run before the neighbour kernels");
    HIPCHK(hipMemsetAsync(h->dep, 0, (size_t)h->N * sizeof(int64_t), h->stream));
    do_density_eos(h, dt, 0.0, 1, false, true);
    launch_rain_merge(h->stream, (int)h->N, h->rain, h->dep);
    h->dep_pending = false;
    launch_momentum_kdk(h->L, h->S, h->stream, mom_args(h, dt, 0));
    LAUNCHCHK();
    h->initialized = true;
    return ASPH_OK;
}

int asph_kick_drift(asph_t* h, float dt) {
    if (!h) return fail("null handle");
    do_kick_drift(h, dt, 2, true);
    LAUNCHCHK();
    return ASPH_OK;
}

int asph_density_eos(asph_t* h, float dt, double simt) {
    if (!h) return fail("null handle");
    if (!h->sorted) return fail("asph_sort must run before the neighbour kernels");
    do_density_eos(h, dt, simt, 1 | 2, false, true);
    LAUNCHCHK();
    return ASPH_OK;
}

int asph_momentum_kdk(asph_t* h, float dt) {
    if (!h) return fail("null handle");
    if (!h->sorted) return fail("asph_sort must run before the neighbour kernels");
    do_momentum_kdk(h, dt, true);
    LAUNCHCHK();
    return ASPH_OK;
}

int asph_rain(asph_t* h, float dt) {
    if (!h) return fail("null handle");
    do_rain(h, dt, true);
    LAUNCHCHK();
    return ASPH_OK;
}

int asph_step(asph_t* h, float dt, double simt, int32_t sort) {
    if (!h) return fail("null handle");
    if (!sort && !h->sorted) return fail("asph_sort must run before the neighbour kernels");
    issue_step(h, dt, simt, sort != 0, false, true);
    LAUNCHCHK();
    return ASPH_OK;
}

int asph_max_vel(asph_t* h, float out[3]) {
    if (!h || !out) return fail("null argument");
    return read_maxv(h, 0, h->n_members(), out);
}

int asph_member_max_vel(asph_t* h, int32_t m, float out[3]) {
    if (!h || !out || m < 0 || m >= h->n_members()) return fail("bad argument");
    return read_maxv(h, m, m + 1, out);
}

int asph_clamp_vel(asph_t* h, float vmax) {
    if (!h) return fail("null handle");
    launch_clamp_vel(h->L, h->stream, h->vel, vmax, 0.f);
    LAUNCHCHK();
    return ASPH_OK;
}

int asph_audit(asph_t* h, double out[7]) {
    if (!h || !out) return fail("null argument");
    HIPCHK(hipMemsetAsync(h->audit_buf, 0, 8 * sizeof(double), h->stream));
    launch_audit(h->L, h->stream, h->pos, h->vel, h->thq, h->rain, h->dep_pending ? h->dep : nullptr,
                 h->precip, h->audit_buf);
    LAUNCHCHK();
    double t[8];
    HIPCHK(hipMemcpyAsync(t, h->audit_buf, sizeof(t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < 7; ++k) out[k] = t[k];
    return ASPH_OK;
}

int asph_get_clock(const asph_t* h, double* simt, double* since_sort, float* sort_dt) {
    if (!h) return fail("null handle");
    if (simt) *simt = h->simt;
    if (since_sort) *since_sort = h->since_sort;
    if (sort_dt) *sort_dt = h->sort_dt;
    return ASPH_OK;
}

int asph_coarse_dump(asph_t* h, int32_t sx, int32_t sy, int32_t sz, float* out) {
    if (!h || !out) return fail("null argument");
    if (!h->sorted) return fail("no cell table: run asph_sort first");
    if (sx < 1 || sy < 1 || sz < 1) return fail("dump strides must be >= 1");
    const DevParams& P = h->mem_h[0].P;
    int cx = (P.nx + sx - 1) / sx, cy = (P.ny + sy - 1) / sy, cz = (P.nz + sz - 1) / sz;
    size_t n = (size_t)cx * cy * cz * 6;
    float* d = nullptr;
    HIPCHK(hipMalloc(&d, n * sizeof(float)));
    launch_coarse_dump(h->L, 0, h->stream, h->cstart, h->vel, h->thq, d, sx, sy, sz, cx, cy, cz);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail("coarse dump: %s", hipGetErrorString(e));
    return ASPH_OK;
}

int asph_column_diag(asph_t* h, float* out4) {
    if (!h || !out4) return fail("null argument");
    if (!h->sorted) return fail("no cell table: run asph_sort first");
    const int ncol = h->mem_h[0].P.ncol;
    float4* d = nullptr;
    HIPCHK(hipMalloc(&d, (size_t)ncol * sizeof(float4)));
    launch_column_diag(h->L, 0, ncol, h->stream, h->cstart, h->thq, h->vel, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out4, d, (size_t)ncol * sizeof(float4), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail("column diag: %s", hipGetErrorString(e));
    return ASPH_OK;
}

// DumpHeader, asph_h100.cu:743-747 (68 bytes, no padding).
struct DumpHeader {
    uint32_t magic, version;
    int32_t nx, ny, nz, cx, cy, cz, sx, sy, sz, nfields;
    float t_sim, dt, lx, ly, lz;
};
static_assert(sizeof(DumpHeader) == 68, "dump header must be 68 bytes");

int asph_write_dump(asph_t* h, const char* outdir, int32_t sx, int32_t sy, int32_t sz, double simt,
                    float dt, char* path_out, size_t path_cap) {
    if (!h || !outdir) return fail("null argument");
    const DevParams& P = h->mem_h[0].P;
    int cx = (P.nx + sx - 1) / sx, cy = (P.ny + sy - 1) / sy, cz = (P.nz + sz - 1) / sz;
    std::vector<float> buf((size_t)cx * cy * cz * 6);
    if (asph_coarse_dump(h, sx, sy, sz, buf.data()) != ASPH_OK) return ASPH_ERR;
    mkdir(outdir, 0775);
    char fn[1024];
    snprintf(fn, sizeof(fn), "%s/asph_t%06.0fs.bin", outdir, simt);   // asph_h100.cu:1107
    FILE* f = fopen(fn, "wb");
    if (!f) return fail("cannot open %s", fn);
    DumpHeader hh{};
    hh.magic = 0x41535048u; hh.version = 3;                           // :1111-1116
    hh.nx = P.nx; hh.ny = P.ny; hh.nz = P.nz;
    hh.cx = cx; hh.cy = cy; hh.cz = cz;
    hh.sx = sx; hh.sy = sy; hh.sz = sz;
    hh.nfields = 6; hh.t_sim = (float)simt; hh.dt = dt;
    hh.lx = P.Lx; hh.ly = P.Ly; hh.lz = P.Lz;
    bool ok = fwrite(&hh, sizeof(hh), 1, f) == 1;
    ok = ok && fwrite(buf.data(), sizeof(float), buf.size(), f) == buf.size();
    fclose(f);
    if (!ok) return fail("short write to %s", fn);
    if (path_out && path_cap) snprintf(path_out, path_cap, "%s", fn);
    return ASPH_OK;
}

int asph_run_hooked(asph_t* h, const asph_run_args* a, int32_t report_every, asph_run_hook hook, void* user,
                    asph_run_stats* st) {
    if (!h || !a) return fail("null argument");
    if (h->n_members() > 1 && !a->fixed_dt) return fail("an ensemble runs with a fixed dt");
    if (a->sim_end <= 0 && a->max_steps <= 0 && !hook)
        return fail("asph_run needs sim_end or max_steps (the loop would not end)");
    const asph_config& c = h->cfgs[0];
    const float D = std::fmin(std::fmin(c.dx, c.dy), c.dz);
    const float dt_static = c.cfl * D / c.cs;                                          // :794
    const float dt_cap = c.dt_cap > 0.f ? std::fmin(c.dt_cap, dt_static) : dt_static;  // :795
    const float dt_floor = 0.05f * dt_cap;                                             // :796
    float dt = a->dt > 0.f ? a->dt : dt_cap;
    float dt_required = dt;
    if (!h->initialized) {                                                             // :899-921
        if (asph_sort(h) != ASPH_OK) return ASPH_ERR;
        if (asph_force(h, dt) != ASPH_OK) return ASPH_ERR;
    }
    double next_dump = a->dump_every > 0 ? (std::floor(h->simt / a->dump_every) + 1) * a->dump_every : 1e300;
    int64_t step = 0, sorts = 0, trips = 0, dumps = 0;
    float u[3] = {0, 0, 0};
    int rc = ASPH_OK;
    auto t0 = std::chrono::steady_clock::now();
    // fills and delivers an event to the hook; nonzero from the hook stops the run
    auto emit = [&](int32_t kind, const char* path) -> int {
        if (!hook) return 0;
        asph_run_event ev;
        ev.kind = kind;
        ev.step = step; ev.sorts = sorts; ev.trips = trips; ev.dumps = dumps;
        ev.simt = h->simt;
        ev.wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        ev.dt = dt; ev.dt_required = dt_required;
        ev.umax[0] = u[0]; ev.umax[1] = u[1]; ev.umax[2] = u[2];
        ev.dump_path = path;
        return hook(user, h, &ev);
    };
    while ((a->sim_end <= 0 || h->simt < a->sim_end) && (a->max_steps <= 0 || step < a->max_steps)) {
        ++step;
        bool sort = (c.flags & ASPH_F_SORT_EVERY_STEP) || h->since_sort >= h->sort_dt;  // :963
        if (sort) { h->since_sort = 0.0; ++sorts; }
        if (asph_step(h, dt, h->simt, sort ? 1 : 0) != ASPH_OK) return ASPH_ERR;
        h->simt += dt; h->since_sort += dt;                                             // :1038
        if (asph_max_vel(h, u) != ASPH_OK) return ASPH_ERR;                             // :1043-1046
        dt_required = c.cfl * std::fmin(std::fmin(c.dx / (c.cs + u[0]), c.dy / (c.cs + u[1])),
                                        c.dz / (c.cs + u[2])) * 0.90f;
        float adopt = a->fixed_dt ? dt : std::fmin(dt_cap, dt_required);
        if (!a->fixed_dt) {
            if (adopt < dt_floor) {                                                     // :1054-1071
                ++trips;
                if (a->verbose)   // asph_h100.cu:1057-1061
                    fprintf(stderr,
                            "[CFL] step %lld t=%.1fs : required %.4f s < floor %.4f s "
                            "(|u|=%.1f |v|=%.1f |w|=%.1f; binding %s)\n",
                            (long long)step, h->simt, adopt, dt_floor, u[0], u[1], u[2],
                            (c.dz / (c.cs + u[2]) <= std::fmin(c.dx / (c.cs + u[0]), c.dy / (c.cs + u[1]))) ? "z" : "xy");
                if (c.strict_cfl || trips > c.max_trip) {
                    if (a->verbose)
                        fprintf(stderr, "[CFL] ABORT after %lld violation(s). The scheme is outside its stability "
                                        "envelope.\n", (long long)trips);
                    rc = ASPH_CFL_ABORT;
                    break;
                }
                launch_clamp_vel(h->L, h->stream, h->vel, 0.f, 0.5f);
                LAUNCHCHK();
                adopt = std::fmax(dt_floor, c.cfl * D / (1.5f * c.cs) * 0.90f);
            } else if (adopt < 0.98f * dt) {
                trips = 0;                                                              // :1072
            }
            dt = adopt;
        }
        h->sort_dt = sort_interval(c, u, dt);                                           // :1075-1077
        if (h->simt >= next_dump) {                                                     // :1081-1147
            char path[1024];
            path[0] = 0;
            if (a->outdir) {
                if (asph_write_dump(h, a->outdir, a->dump_sx, a->dump_sy, a->dump_sz, h->simt, dt,
                                    path, sizeof(path)) != ASPH_OK)
                    return ASPH_ERR;
                ++dumps;
                if (a->verbose && !hook) printf("%-8lld +%-8.0fs dt %.4f  CFL %.2f  -> %s\n", (long long)step,
                                                h->simt, dt, dt_required > 0 ? dt / dt_required : 0.f, path);
            }
            if (a->validate && !hook) {
                double tt[7];
                if (asph_audit(h, tt) != ASPH_OK) return ASPH_ERR;
                if (a->verbose)
                    printf("  AUDIT  KE %.4e  PE %.4e  IE %.4e  E_tot %.4e  water %.4e  rain %.4e  pz %.4e\n",
                           tt[0], tt[1], tt[2], tt[0] + tt[1] + tt[2], tt[3], tt[6], tt[4]);
            }
            next_dump += a->dump_every;
            if (emit(ASPH_EV_DUMP, path[0] ? path : nullptr)) break;
        }
        if (report_every > 0 && step % report_every == 0 && emit(ASPH_EV_REPORT, nullptr)) break;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (st) {
        st->steps = step; st->sorts = sorts; st->trips = trips; st->dumps = dumps;
        st->simt = h->simt; st->dt = dt; st->dt_required = dt_required;
        st->umax[0] = u[0]; st->umax[1] = u[1]; st->umax[2] = u[2];
        st->wall_s = wall;
    }
    return rc;
}

int asph_run(asph_t* h, const asph_run_args* a, asph_run_stats* st) {
    return asph_run_hooked(h, a, 0, nullptr, nullptr, st);
}

int asph_run_device(asph_t* h, const asph_run_args* a, int32_t poll, asph_run_stats* st) {
    if (!h || !a) return fail("null argument");
    if (h->n_members() > 1) return fail("the device controller runs single-domain handles");
    if (a->sim_end <= 0 && a->max_steps <= 0) return fail("asph_run_device needs sim_end or max_steps");
    const asph_config& c = h->cfgs[0];
    const float D = std::fmin(std::fmin(c.dx, c.dy), c.dz);
    const float dt_static = c.cfl * D / c.cs;                                          // :794
    const float dt_cap = c.dt_cap > 0.f ? std::fmin(c.dt_cap, dt_static) : dt_static;  // :795
    const float dt_floor = 0.05f * dt_cap;                                             // :796
    const float dt0 = a->dt > 0.f ? a->dt : dt_cap;
    if (!h->initialized) {
        if (asph_sort(h) != ASPH_OK) return ASPH_ERR;
        if (asph_force(h, dt0) != ASPH_OK) return ASPH_ERR;
    }
    const int pairs = std::max(1, (poll + 1) / 2);   // replays of the two-step graph between polls
    if (!h->ctl_d && h->alloc(&h->ctl_d, 1) != 0) return fail("out of device memory (controller)");
    StepCtl k;
    std::memset(&k, 0, sizeof(k));
    k.simt = h->simt; k.since_sort = h->since_sort; k.sim_end = a->sim_end;
    k.dump_every = a->dump_every > 0 ? a->dump_every : 0.0;
    k.next_dump = k.dump_every > 0 ? (std::floor(h->simt / k.dump_every) + 1) * k.dump_every : 0.0;
    k.max_steps = a->max_steps; k.n = h->N;
    k.dt = dt0; k.dt_required = dt0; k.sort_dt = h->sort_dt;
    k.clamp_dt = std::fmax(dt_floor, c.cfl * D / (1.5f * c.cs) * 0.90f);               // :1068-1069
    k.dt_cap = dt_cap; k.dt_floor = dt_floor; k.cfl = c.cfl; k.cs = c.cs; k.dx = c.dx; k.dy = c.dy; k.dz = c.dz;
    k.max_trip = c.max_trip; k.strict = c.strict_cfl; k.every = (c.flags & ASPH_F_SORT_EVERY_STEP) ? 1 : 0;
    k.fixed = a->fixed_dt ? 1 : 0;
    HIPCHK(hipMemcpyAsync(h->ctl_d, &k, sizeof(k), hipMemcpyHostToDevice, h->stream));
    h->L.ctl = h->ctl_d;
    auto t0 = std::chrono::steady_clock::now();
    int rc = ASPH_OK;
    int64_t dumps = 0;
    // one step: controller, sort stages (gated), the four stages, controller, clamp (gated)
    auto device_step = [&](bool launch) {
        if (launch) launch_ctl_pre(h->stream, h->ctl_d);
        do_sort(h, launch);
        do_kick_drift(h, dt0, 2, launch);
        do_density_eos(h, dt0, 0.0, 1 | 2, true, launch);
        do_momentum_kdk(h, dt0, launch);
        do_rain(h, dt0, launch);
        if (launch) {
            launch_ctl_post(h->stream, h->ctl_d, h->maxv);
            launch_clamp_vel_ctl(h->L, h->stream, h->vel);
        }
    };
    for (;;) {
        for (int r = 0; r < pairs; ++r) {
            // a captured pair of steps returns the rotated buffers to their arrangement
            auto key = std::make_tuple((void*)h->pos, (void*)h->vel, (void*)h->thq, (void*)h->rho,
                                       (void*)h->acc, (void*)h->rain, 3, 0.f);
            auto it = h->graphs.find(key);
            if (it == h->graphs.end()) {
                hipGraph_t g = nullptr;
                hipGraphExec_t ge = nullptr;
                g_capturing = true;
                HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                device_step(true);
                device_step(true);
                hipError_t e = hipStreamEndCapture(h->stream, &g);
                g_capturing = false;
                if (e != hipSuccess) { h->L.ctl = nullptr; return fail("graph capture: %s", hipGetErrorString(e)); }
                HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
                (void)hipGraphDestroy(g);
                h->graphs[key] = ge;
                HIPCHK(hipGraphLaunch(ge, h->stream));
            } else {
                HIPCHK(hipGraphLaunch(it->second, h->stream));
                device_step(false);
                device_step(false);
            }
        }
        HIPCHK(hipMemcpyAsync(&k, h->ctl_d, sizeof(k), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (k.abort) {
            if (a->verbose)
                fprintf(stderr, "[CFL] ABORT after %lld violation(s) at step %lld. The scheme is outside its "
                                "stability envelope.\n", (long long)k.trips, (long long)k.steps);
            rc = ASPH_CFL_ABORT;
            break;
        }
        if (k.pause) {   // a dump time was reached: the state is that of the step that reached it
            h->simt = k.simt;
            if (a->outdir) {
                char path[1024];
                if (asph_write_dump(h, a->outdir, a->dump_sx, a->dump_sy, a->dump_sz, k.simt, k.dt, path,
                                    sizeof(path)) != ASPH_OK) {
                    h->L.ctl = nullptr;
                    return ASPH_ERR;
                }
                ++dumps;
                if (a->verbose) printf("%-8lld +%-8.0fs dt %.4f  CFL %.2f  -> %s\n", (long long)k.steps, k.simt,
                                       k.dt, k.dt_required > 0 ? k.dt / k.dt_required : 0.f, path);
            }
            if (a->validate) {
                double tt[7];
                if (asph_audit(h, tt) != ASPH_OK) { h->L.ctl = nullptr; return ASPH_ERR; }
                if (a->verbose)
                    printf("  AUDIT  KE %.4e  PE %.4e  IE %.4e  E_tot %.4e  water %.4e  rain %.4e  pz %.4e\n",
                           tt[0], tt[1], tt[2], tt[0] + tt[1] + tt[2], tt[3], tt[6], tt[4]);
            }
            k.pause = 0;
            k.next_dump += k.dump_every;
            HIPCHK(hipMemcpyAsync(h->ctl_d, &k, sizeof(k), hipMemcpyHostToDevice, h->stream));
            continue;
        }
        if (k.idle) break;   // the run's end was reached
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    h->L.ctl = nullptr;
    h->simt = k.simt;
    h->since_sort = k.since_sort;
    h->sort_dt = k.sort_dt;
    if (st) {
        st->steps = k.steps; st->sorts = k.sorts; st->trips = k.trips; st->dumps = dumps;
        st->simt = k.simt; st->dt = k.dt; st->dt_required = k.dt_required;
        st->umax[0] = k.u[0]; st->umax[1] = k.u[1]; st->umax[2] = k.u[2];
        st->wall_s = wall;
    }
    return rc;
}

int asph_run_graph(asph_t* h, int64_t nsteps, float dt, int32_t sort_every, asph_run_stats* st) {
    if (!h) return fail("null handle");
    if (nsteps < 0 || !(dt > 0.f)) return fail("bad steps or dt");
    const bool every = (h->cfgs[0].flags & ASPH_F_SORT_EVERY_STEP) != 0;
    if (!h->initialized) {
        if (asph_sort(h) != ASPH_OK) return ASPH_ERR;
        if (asph_force(h, dt) != ASPH_OK) return ASPH_ERR;
        h->since_sort = 1e9;
    }
    // the device clock holds the start time of the step about to run; each replay advances it
    launch_step_begin(h->stream, h->disp, h->disp_slots, 0, h->clock, 0, h->simt - (double)dt, 0.f);
    LAUNCHCHK();
    int64_t sorts = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int64_t k = 0; k < nsteps; ++k) {
        bool sort = every || (sort_every > 0 ? (k % sort_every == 0) : h->since_sort >= h->sort_dt);
        auto key = std::make_tuple((void*)h->pos, (void*)h->vel, (void*)h->thq, (void*)h->rho,
                                   (void*)h->acc, (void*)h->rain, sort ? 1 : 0, dt);
        auto it = h->graphs.find(key);
        if (it == h->graphs.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            g_capturing = true;
            HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            issue_step(h, dt, 0.0, sort, true, true);
            hipError_t e = hipStreamEndCapture(h->stream, &g);
            g_capturing = false;
            if (e != hipSuccess) return fail("graph capture: %s", hipGetErrorString(e));
            HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            (void)hipGraphDestroy(g);
            h->graphs[key] = ge;
            HIPCHK(hipGraphLaunch(ge, h->stream));
        } else {
            HIPCHK(hipGraphLaunch(it->second, h->stream));
            issue_step(h, dt, 0.0, sort, true, false);
        }
        if (sort) { ++sorts; h->since_sort = 0.0; }
        h->since_sort += dt;
        h->simt += dt;
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    float u[3] = {0, 0, 0};
    if (nsteps > 0 && read_maxv(h, 0, h->n_members(), u) != ASPH_OK) return ASPH_ERR;
    if (nsteps > 0) h->sort_dt = sort_interval(h->cfgs[0], u, dt);
    if (st) {
        st->steps = nsteps; st->sorts = sorts; st->trips = 0; st->dumps = 0;
        st->simt = h->simt; st->dt = dt; st->dt_required = 0.f;
        st->umax[0] = u[0]; st->umax[1] = u[1]; st->umax[2] = u[2];
        st->wall_s = wall;
    }
    return ASPH_OK;
}

int asph_set_profiling(asph_t* h, int32_t enable) {
    if (!h) return fail("null handle");
    h->prof = enable != 0;
    return ASPH_OK;
}

int asph_kernel_times(asph_t* h, double ms[ASPH_K_COUNT], int64_t counts[ASPH_K_COUNT]) {
    if (!h) return fail("null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < ASPH_K_COUNT; ++k) {
        double tot = 0;
        for (auto& pr : h->ev_live[k]) {
            float t = 0.f;
            if (hipEventElapsedTime(&t, pr.first, pr.second) == hipSuccess) tot += t;
            h->ev_pool.push_back(pr.first);
            h->ev_pool.push_back(pr.second);
        }
        if (ms) ms[k] = tot;
        if (counts) counts[k] = (int64_t)h->ev_live[k].size();
        h->ev_live[k].clear();
    }
    return ASPH_OK;
}

// --- tuning of the brick decomposition --------------------------------------------------------
}  // extern "C"

namespace {

void drop_graphs(asph_t* h) {
    (void)hipStreamSynchronize(h->stream);
    for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
    h->graphs.clear();
}

// Density without the EOS/microphysics epilogue into the scratch density buffer: reads the
// state, writes nothing the step keeps.
DensArgs trial_dens_args(asph_t* h, float dt) {
    DensArgs A;
    A.pos = h->pos; A.rho_old = h->rho; A.rho_new = h->rho2; A.thq = h->thq; A.prs = h->prs;
    A.vel = h->vel; A.cstart = h->cstart; A.keys_sorted = h->keys_sorted; A.rain = h->rain;
    A.dep = h->dep; A.maxv = h->maxv; A.disp = h->disp; A.disp_slots = h->disp_slots;
    A.bhalo = h->L.bhalo;
    A.clock = nullptr; A.simt = h->simt; A.stats = nullptr; A.s12_out = nullptr; A.rinv = h->rinv;
    A.dt = dt; A.bits = 0;
    return A;
}

}  // namespace

extern "C" {

int asph_set_tuning(asph_t* h, const asph_tuning* t) {
    if (!h || !t) return fail("null argument");
    if (t->chunk < brick_chunk_min() || t->chunk > brick_chunk_max())
        return fail("chunk %d outside [%d, %d]", t->chunk, brick_chunk_min(), brick_chunk_max());
    if (t->order != 0 && t->order != 1) return fail("order must be 0 or 1");
    if (t->chunk != h->L.chunk || t->order != h->L.order) drop_graphs(h);
    h->L.chunk = t->chunk;
    h->L.order = t->order;
    if (h->sorted) {
        launch_brick_items(h->L, h->stream, h->cstart);
        LAUNCHCHK();
    }
    return ASPH_OK;
}

int asph_get_tuning(asph_t* h, asph_tuning* t) {
    if (!h || !t) return fail("null argument");
    t->chunk = h->L.chunk;
    t->order = h->L.order;
    return ASPH_OK;
}

int asph_autotune(asph_t* h, int32_t reps, asph_tune_report* rep) {
    if (!h) return fail("null handle");
    if (!h->sorted || !h->initialized) return fail("autotune needs a sorted, force-evaluated state");
    if (reps < 1) reps = 1;
    static const int32_t chunks[] = {256, 384, 512, 768, 1024};
    std::vector<asph_tuning> cand;
    for (int32_t o = 1; o >= 0; --o)
        for (int32_t c : chunks) cand.push_back({c, o});
    const size_t N = (size_t)h->N;
    const size_t nmv = (size_t)192 * h->n_members();
    float4* acc_save = nullptr;
    float* dth_save = nullptr;
    uint32_t* maxv_save = nullptr;
    if (hipMalloc(&acc_save, N * sizeof(float4)) != hipSuccess) return fail("autotune: out of memory");
    if (hipMalloc(&dth_save, N * sizeof(float)) != hipSuccess ||
        hipMalloc(&maxv_save, nmv * sizeof(uint32_t)) != hipSuccess) {
        (void)hipFree(acc_save);
        (void)hipFree(dth_save);
        return fail("autotune: out of memory");
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const asph_tuning keep = {h->L.chunk, h->L.order};
    drop_graphs(h);
    int rc = ASPH_OK;
    std::vector<double> ms(cand.size(), 0.0);
    std::vector<uint32_t> nit(cand.size(), 0u);
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 256;
    const float dt = 0.1f;   // enters only the skipped epilogue
    hipError_t e = hipMemcpyAsync(acc_save, h->acc, N * sizeof(float4), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dth_save, h->dthdt, N * sizeof(float), hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(maxv_save, h->maxv, nmv * sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream);
    for (size_t k = 0; k < cand.size() && e == hipSuccess; ++k) {
        h->L.chunk = cand[k].chunk;
        h->L.order = cand[k].order;
        launch_brick_items(h->L, h->stream, h->cstart);
        for (int r = -1; r < reps; ++r) {   // r = -1: warm-up
            if (r == 0) (void)hipEventRecord(e0, h->stream);
            launch_density_eos(h->L, h->S, h->stream, trial_dens_args(h, dt));
            launch_momentum_kdk(h->L, h->S, h->stream, mom_args(h, dt, 0));
        }
        (void)hipEventRecord(e1, h->stream);
        e = hipEventSynchronize(e1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess) e = hipMemcpy(&nit[k], h->L.nitems, sizeof(uint32_t), hipMemcpyDeviceToHost);
        ms[k] = (double)t / reps;
    }
    // The heaviest-first order unless brick order is more than 10% faster than every
    // heaviest-first shape; within the chosen order the fastest shape, except that among the
    // heaviest-first shapes within 3% of their fastest the largest chunk that still gives two
    // workgroups per CU is kept (or, when none does, the smallest chunk). On a near-uniform state
    // the shapes tie: in a large domain, large chunks and the heaviest-first order then protect
    // the launch from the dense regions that develop later (the long workgroups start first, and
    // fewer items restage a dense brick's halo; brick order, chosen at t = 0 of the 128^3 run for
    // being more than 3% faster there, made its 600 steps 29% slower); in a small one, a launch
    // is as long as one workgroup's chain, and small chunks shorten it.
    size_t best = 0, best1 = cand.size();
    for (size_t k = 1; k < cand.size(); ++k)
        if (ms[k] < ms[best]) best = k;
    for (size_t k = 0; k < cand.size(); ++k)
        if (cand[k].order == 1 && (best1 == cand.size() || ms[k] < ms[best1])) best1 = k;
    if (best1 < cand.size() && ms[best1] <= 1.10 * ms[best]) best = best1;
    size_t keep1 = cand.size(), small1 = cand.size();
    for (size_t k = 0; k < cand.size() && cand[best].order == 1; ++k) {
        if (cand[k].order != 1 || ms[k] > 1.03 * ms[best]) continue;
        if (nit[k] >= 2u * (uint32_t)ncu && (keep1 == cand.size() || cand[k].chunk > cand[keep1].chunk)) keep1 = k;
        if (small1 == cand.size() || cand[k].chunk < cand[small1].chunk) small1 = k;
    }
    if (keep1 == cand.size()) keep1 = small1;
    if (keep1 < cand.size()) best = keep1;
    if (e != hipSuccess) {
        rc = fail("autotune: %s", hipGetErrorString(e));
        h->L.chunk = keep.chunk;
        h->L.order = keep.order;
    } else {
        h->L.chunk = cand[best].chunk;
        h->L.order = cand[best].order;
    }
    // restore the forces of the state and the items of the chosen decomposition
    (void)hipMemcpyAsync(h->acc, acc_save, N * sizeof(float4), hipMemcpyDeviceToDevice, h->stream);
    (void)hipMemcpyAsync(h->dthdt, dth_save, N * sizeof(float), hipMemcpyDeviceToDevice, h->stream);
    (void)hipMemcpyAsync(h->maxv, maxv_save, nmv * sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream);
    launch_brick_items(h->L, h->stream, h->cstart);
    (void)hipStreamSynchronize(h->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(acc_save);
    (void)hipFree(dth_save);
    (void)hipFree(maxv_save);
    if (rep) {
        rep->n = (int32_t)std::min<size_t>(cand.size(), ASPH_TUNE_MAX);
        for (int32_t k = 0; k < rep->n; ++k) {
            rep->cand[k] = cand[k];
            rep->ms[k] = ms[k];
            rep->items[k] = nit[k];
        }
        rep->best = (int32_t)best;
        rep->cus = ncu;
    }
    if (rc == ASPH_OK) LAUNCHCHK();
    return rc;
}

int asph_set_counters(asph_t* h, int32_t enable) {
    if (!h) return fail("null handle");
    if (enable && !h->stats_buf) {
        int rc = h->alloc(&h->stats_buf, (size_t)ASPH_C_COUNT * 64);   // 64 slots per counter
        if (rc) return rc;
    }
    unsigned long long* want = enable ? h->stats_buf : nullptr;
    if (want != h->stats) {
        // captured graphs hold the counter pointer as a kernel argument
        HIPCHK(hipStreamSynchronize(h->stream));
        for (auto& kv : h->graphs) (void)hipGraphExecDestroy(kv.second);
        h->graphs.clear();
    }
    if (enable)
        HIPCHK(hipMemsetAsync(h->stats_buf, 0, ASPH_C_COUNT * 64 * sizeof(unsigned long long), h->stream));
    h->stats = want;
    return ASPH_OK;
}

int asph_get_counters(asph_t* h, uint64_t out[ASPH_C_COUNT]) {
    if (!h || !out) return fail("null argument");
    if (!h->stats_buf) {
        for (int k = 0; k < ASPH_C_COUNT; ++k) out[k] = 0;
        return ASPH_OK;
    }
    std::vector<uint64_t> slots((size_t)ASPH_C_COUNT * 64);
    HIPCHK(hipMemcpyAsync(slots.data(), h->stats_buf, slots.size() * sizeof(uint64_t),
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemsetAsync(h->stats_buf, 0, slots.size() * sizeof(uint64_t), h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < ASPH_C_COUNT; ++k) {
        uint64_t t = 0;
        for (int j = 0; j < 64; ++j) t += slots[(size_t)k * 64 + j];
        out[k] = t;
    }
    return ASPH_OK;
}

int asph_synchronize(asph_t* h) {
    if (!h) return fail("null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return ASPH_OK;
}

}  // extern "C"
