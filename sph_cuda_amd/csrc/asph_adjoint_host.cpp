// asph_adjoint_host.cpp - the adjoint driver behind asph_adjoint_rain: forward window with
// checkpoints, then segment by segment (last first) the reverse sweep of asph_adjoint.hip on a
// tape recorded by recomputing the segment from its checkpoint.
//
// Forward schedule: the state is sorted and the forces evaluated (asph_force, asph_h100.cu:913-920),
// then every step follows the reference loop (asph_h100.cu:962-1038) with the neighbour structure
// rebuilt at the start of every segment of `checkpoint` steps and kept in between (the stale
// cell lists of the reference between gated sorts). A checkpoint is the state at a segment
// start before its sort; recomputing a segment replays the sort, so keys, permutation and cell
// table are those of the forward. The last segment is taped while the forward runs it, so only
// the others are recomputed. Tape records are batched copies; inside a segment a step's start
// position, acceleration and density are the previous step's recorded values.
#include <chrono>
#include <cstring>
#include <vector>

#include "asph_handle.h"

using namespace asph_host;

namespace {

struct Alloc {
    std::vector<void*> ptrs;
    std::vector<hipEvent_t> events;
    hipStream_t stream = nullptr;
    template <class T> T* get(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
    hipEvent_t event() {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        events.push_back(e);
        return e;
    }
    ~Alloc() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
        for (void* p : ptrs) (void)hipFree(p);
    }
};

// The neighbour structure of a segment's sort, kept with its tape: the reverse reads it while the
// handle already sorts for the next recomputation.
struct SegSort {
    uint32_t *keys, *cstart, *perm, *nitems;
    int2* items;
};

struct Checkpoint {
    float4 *pos, *vel, *acc, *thq;
    float* rho;
    int64_t *rain, *precip;
    double simt;
};

struct TapeStep {
    float4 *x0, *v0, *a0, *t0, *x1, *v1, *t1, *a1, *av;
    float *r0, *r1, *delta;
    float2* s12;
    int64_t* rain0;
    // x0, a0 and r0 as used: the previous step's x1, a1 and r1 inside a segment (no sort between)
    const float4 *x0v = nullptr, *a0v = nullptr;
    const float* r0v = nullptr;
    double simt = 0.0;   // simulated time at the start of the step
};

#define ACHK(expr)                                                                      \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) return fail("%s: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

int save(asph_t* h, const Checkpoint& c) {
    const size_t N = (size_t)h->N;
    hipStream_t s = h->stream;
    ACHK(hipMemcpyAsync(c.pos, h->pos, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(c.vel, h->vel, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(c.acc, h->acc, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(c.thq, h->thq, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(c.rho, h->rho, N * sizeof(float), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(c.rain, h->rain, N * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(c.precip, h->precip, (size_t)h->ncol * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    return ASPH_OK;
}

int restore(asph_t* h, const Checkpoint& c) {
    const size_t N = (size_t)h->N;
    hipStream_t s = h->stream;
    ACHK(hipMemcpyAsync(h->pos, c.pos, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(h->vel, c.vel, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(h->acc, c.acc, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(h->thq, c.thq, N * sizeof(float4), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(h->rho, c.rho, N * sizeof(float), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(h->rain, c.rain, N * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemcpyAsync(h->precip, c.precip, (size_t)h->ncol * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    ACHK(hipMemsetAsync(h->dep, 0, N * sizeof(int64_t), s));
    h->simt = c.simt;
    h->sorted = false;
    h->dep_pending = false;
    return ASPH_OK;
}

// The force evaluation at t = 0 (asph_force) with its tape.
int taped_force(asph_t* h, TapeStep& t, float dt) {
    const size_t N = (size_t)h->N;
    hipStream_t s = h->stream;
    AdjCopies pre;
    pre.add(t.x0, h->pos, N); pre.add(t.v0, h->vel, N); pre.add(t.a0, h->acc, N); pre.add(t.t0, h->thq, N);
    pre.add(t.r0, h->rho, N); pre.add(t.rain0, h->rain, N);
    launch_adj_copies(s, pre);
    t.x0v = t.x0; t.a0v = t.a0; t.r0v = t.r0;
    t.simt = 0.0;
    ACHK(hipMemsetAsync(h->dep, 0, N * sizeof(int64_t), s));
    h->rec_s12 = t.s12;
    do_density_eos(h, dt, 0.0, 1, false, true);
    h->rec_s12 = nullptr;
    AdjCopies mid;
    mid.add(t.x1, h->pos, N); mid.add(t.v1, h->vel, N); mid.add(t.r1, h->rho, N); mid.add(t.t1, h->thq, N);
    launch_adj_copies(s, mid);
    launch_rain_merge(s, (int)N, h->rain, h->dep);
    h->dep_pending = false;
    h->rec_av = t.av;
    launch_momentum_kdk(h->L, h->S, s, mom_args(h, dt, 0));
    h->rec_av = nullptr;
    launch_adj_delta(s, h->disp, h->disp_slots, t.delta);
    AdjCopies post;
    post.add(t.a1, h->acc, N);
    launch_adj_copies(s, post);
    h->initialized = true;
    ACHK(hipGetLastError());
    return ASPH_OK;
}

// One step of the reference loop (issue_step without a sort) with its tape. prev: the previous
// step of the same segment, whose x1, a1 and r1 are this step's x0, a0 and r0; null after a sort.
int taped_step(asph_t* h, TapeStep& t, const TapeStep* prev, float dt) {
    const size_t N = (size_t)h->N;
    hipStream_t s = h->stream;
    AdjCopies pre;
    pre.add(t.v0, h->vel, N); pre.add(t.t0, h->thq, N); pre.add(t.rain0, h->rain, N);
    if (prev) {
        t.x0v = prev->x1; t.a0v = prev->a1; t.r0v = prev->r1;
    } else {
        pre.add(t.x0, h->pos, N); pre.add(t.a0, h->acc, N); pre.add(t.r0, h->rho, N);
        t.x0v = t.x0; t.a0v = t.a0; t.r0v = t.r0;
    }
    launch_adj_copies(s, pre);
    t.simt = h->simt;
    do_kick_drift(h, dt, 2, true);
    h->rec_s12 = t.s12;
    do_density_eos(h, dt, h->simt, 1 | 2, false, true);
    h->rec_s12 = nullptr;
    AdjCopies mid;
    mid.add(t.x1, h->pos, N); mid.add(t.v1, h->vel, N); mid.add(t.r1, h->rho, N); mid.add(t.t1, h->thq, N);
    launch_adj_copies(s, mid);
    h->rec_av = t.av;
    do_momentum_kdk(h, dt, true);
    h->rec_av = nullptr;
    launch_adj_delta(s, h->disp, h->disp_slots, t.delta);
    AdjCopies post;
    post.add(t.a1, h->acc, N);
    launch_adj_copies(s, post);
    do_rain(h, dt, true);
    h->simt += dt;
    ACHK(hipGetLastError());
    return ASPH_OK;
}

AdjTape tape_view(const SegSort& g, const TapeStep& t) {
    AdjTape T;
    T.x0 = t.x0v; T.v0 = t.v0; T.a0 = t.a0v; T.t0 = t.t0; T.r0 = t.r0v; T.rain0 = t.rain0;
    T.x1 = t.x1; T.v1 = t.v1; T.r1 = t.r1; T.s12 = t.s12; T.t1 = t.t1; T.a1 = t.a1; T.av = t.av;
    T.keys = g.keys; T.cstart = g.cstart; T.delta = t.delta;
    return T;
}

}  // namespace

extern "C" {

int asph_adjoint_rain(asph_t* h, const asph_adjoint_args* a, float* dJ_dtheta0, float* dJ_dthq4,
                      asph_adjoint_result* r) {
    if (!h || !a || !r) return fail("null argument");
    if (h->n_members() != 1) return fail("the adjoint runs on a single-domain handle");
    const asph_config& cfg = h->cfgs[0];
    const DevParams& Ph = h->mem_h[0].P;
    const int S = h->S;
    if (cfg.nx < 2 * S + 3 || cfg.ny < 2 * S + 3)
        return fail("the adjoint needs nx, ny >= 2*stencil+3 (got %d, %d)", cfg.nx, cfg.ny);
    if (h->N >= (int64_t)1 << 30) return fail("the adjoint handles fewer than 2^30 particles");
    const int64_t nsteps = a->nsteps;
    const float dt = a->dt;
    const int ck = a->checkpoint > 0 ? a->checkpoint : 10;
    if (nsteps < 1 || !(dt > 0.f)) return fail("bad nsteps or dt");
    const int64_t nseg = (nsteps + ck - 1) / ck;
    const size_t N = (size_t)h->N;
    hipStream_t s = h->stream;
    const DevParams* Pd = &h->mem_d[0].P;

    Alloc A;
    std::vector<Checkpoint> cps((size_t)nseg);
    for (auto& c : cps) {
        c.pos = A.get<float4>(N); c.vel = A.get<float4>(N); c.acc = A.get<float4>(N);
        c.thq = A.get<float4>(N); c.rho = A.get<float>(N); c.rain = A.get<int64_t>(N);
        c.precip = A.get<int64_t>((size_t)h->ncol);
        if (!c.pos || !c.vel || !c.acc || !c.thq || !c.rho || !c.rain || !c.precip)
            return fail("adjoint: out of memory (checkpoints)");
    }
    // two tape sets: segment sg uses set sg % 2, so the reverse of one segment (stream s1) runs
    // while the handle recomputes the one before it (the handle's stream)
    std::vector<TapeStep> tape[2];
    SegSort ss[2];
    const int nitem_max = h->L.nitem_max;
    for (int b = 0; b < 2; ++b) {
        tape[b].resize((size_t)ck + 1);   // the last slot holds the force evaluation (set 0)
        for (auto& t : tape[b]) {
            t.x0 = A.get<float4>(N); t.v0 = A.get<float4>(N); t.a0 = A.get<float4>(N); t.t0 = A.get<float4>(N);
            t.x1 = A.get<float4>(N); t.v1 = A.get<float4>(N); t.t1 = A.get<float4>(N); t.a1 = A.get<float4>(N);
            t.av = A.get<float4>(N); t.r0 = A.get<float>(N); t.r1 = A.get<float>(N); t.s12 = A.get<float2>(N);
            t.rain0 = A.get<int64_t>(N); t.delta = A.get<float>(4);
            if (!t.delta || !t.x0 || !t.v0 || !t.a0 || !t.t0 || !t.x1 || !t.v1 || !t.t1 || !t.a1 || !t.av || !t.r0 ||
                !t.r1 || !t.s12 || !t.rain0)
                return fail("adjoint: out of memory (tape)");
        }
        ss[b].keys = A.get<uint32_t>(N); ss[b].perm = A.get<uint32_t>(N);
        ss[b].cstart = A.get<uint32_t>((size_t)h->nkeys + 1); ss[b].items = A.get<int2>((size_t)nitem_max);
        ss[b].nitems = A.get<uint32_t>(1);
        if (!ss[b].keys || !ss[b].perm || !ss[b].cstart || !ss[b].items || !ss[b].nitems)
            return fail("adjoint: out of memory (sort)");
    }
    hipEvent_t e_rec[2] = {A.event(), A.event()}, e_rev[2] = {A.event(), A.event()};
    if (!e_rec[0] || !e_rec[1] || !e_rev[0] || !e_rev[1]) return fail("adjoint: cannot create events");
    ACHK(hipStreamCreateWithFlags(&A.stream, hipStreamNonBlocking));   // default priority
    hipStream_t s1 = A.stream;
    AdjState st[2];
    for (auto& x : st) {
        x.x = A.get<float4>(N); x.v = A.get<float4>(N); x.a = A.get<float4>(N); x.t = A.get<float4>(N);
        x.rho = A.get<float>(N); x.rain = A.get<double>(N);
        if (!x.x || !x.v || !x.a || !x.t || !x.rho || !x.rain) return fail("adjoint: out of memory (state)");
    }
    AdjWork W;
    W.Ab = A.get<float4>(N); W.AVb = A.get<float4>(N); W.x1b = A.get<float4>(N); W.v1b = A.get<float4>(N);
    W.t2b = A.get<float4>(N); W.pb = A.get<float>(N); W.r1b = A.get<float>(N); W.sb = A.get<float2>(N);
    W.rb_cur = A.get<double>(N);
    double* qc0b = A.get<double>(1);
    if (!W.Ab || !W.AVb || !W.x1b || !W.v1b || !W.t2b || !W.pb || !W.r1b || !W.sb || !W.rb_cur || !qc0b)
        return fail("adjoint: out of memory (work)");

    // ---- forward window; the last segment is taped as it runs ----------------------------------
    if (reset_aux(h, true) != ASPH_OK) return ASPH_ERR;   // as after asph_set_state
    h->rec_s12 = nullptr;
    h->rec_av = nullptr;
    ACHK(hipStreamSynchronize(s));
    auto t0 = std::chrono::steady_clock::now();
    TapeStep& tinit = tape[0][(size_t)ck];
    const int64_t klast = (nseg - 1) * ck;   // first step of the last segment
    // runs segment sg's steps with the tape (after its sort); the force evaluation first for sg = 0
    auto taped_segment = [&](int64_t sg) -> int {
        const int b = (int)(sg & 1);
        AdjCopies c;
        c.add(ss[b].keys, h->keys_sorted, N); c.add(ss[b].perm, h->perm, N);
        c.add(ss[b].cstart, h->cstart, (size_t)h->nkeys + 1); c.add(ss[b].items, h->L.items, (size_t)nitem_max);
        c.add(ss[b].nitems, h->L.nitems, 1);
        launch_adj_copies(s, c);
        if (sg == 0 && taped_force(h, tinit, dt) != ASPH_OK) return ASPH_ERR;
        std::vector<TapeStep>& tp = tape[b];
        const int64_t k0 = sg * ck, k1 = std::min<int64_t>(nsteps, k0 + ck);
        for (int64_t k = k0; k < k1; ++k)
            if (taped_step(h, tp[(size_t)(k - k0)], k > k0 ? &tp[(size_t)(k - k0 - 1)] : nullptr, dt) != ASPH_OK)
                return ASPH_ERR;
        ACHK(hipEventRecord(e_rec[b], s));
        return ASPH_OK;
    };
    cps[0].simt = 0.0;
    if (save(h, cps[0]) != ASPH_OK) return ASPH_ERR;
    do_sort(h, true);
    if (nseg == 1) {
        if (taped_segment(0) != ASPH_OK) return ASPH_ERR;
    } else {
        if (asph_force(h, dt) != ASPH_OK) return ASPH_ERR;
        for (int64_t k = 0; k < klast; ++k) {
            bool srt = k > 0 && k % ck == 0;
            if (srt) {
                cps[(size_t)(k / ck)].simt = h->simt;
                if (save(h, cps[(size_t)(k / ck)]) != ASPH_OK) return ASPH_ERR;
            }
            issue_step(h, dt, h->simt, srt, false, true);
            h->simt += dt;
        }
        cps[(size_t)(nseg - 1)].simt = h->simt;
        if (save(h, cps[(size_t)(nseg - 1)]) != ASPH_OK) return ASPH_ERR;
        do_sort(h, true);
        if (taped_segment(nseg - 1) != ASPH_OK) return ASPH_ERR;
    }
    ACHK(hipStreamSynchronize(s));
    const double fwd_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    double J = 0.0;
    {
        std::vector<int64_t> rain(N);
        ACHK(hipMemcpy(rain.data(), h->rain, N * sizeof(int64_t), hipMemcpyDeviceToHost));
        long double sum = 0;
        for (int64_t v : rain) sum += (long double)v;
        J = (double)(sum * (long double)FX_INV) * (double)Ph.pmass;
    }

    // ---- reverse: segments last first on stream s1; the segment before is recomputed meanwhile ----
    t0 = std::chrono::steady_clock::now();
    int cur = 0;
    ACHK(hipMemsetAsync(st[cur].x, 0, N * sizeof(float4), s1));
    ACHK(hipMemsetAsync(st[cur].v, 0, N * sizeof(float4), s1));
    ACHK(hipMemsetAsync(st[cur].a, 0, N * sizeof(float4), s1));
    ACHK(hipMemsetAsync(st[cur].t, 0, N * sizeof(float4), s1));
    ACHK(hipMemsetAsync(st[cur].rho, 0, N * sizeof(float), s1));
    ACHK(hipMemsetAsync(qc0b, 0, sizeof(double), s1));
    {
        std::vector<double> r0(N, (double)Ph.pmass);   // dJ/d(rain in particle masses)
        ACHK(hipMemcpyAsync(st[cur].rain, r0.data(), N * sizeof(double), hipMemcpyHostToDevice, s1));
        ACHK(hipStreamSynchronize(s1));
    }
    const float frac = (cfg.flags & ASPH_F_SEDIMENT) ? rain_frac(Ph, dt) : 0.f;
    bool rev_issued[2] = {false, false};
    for (int64_t sg = nseg - 1; sg >= 0; --sg) {
        if (sg > 0) {   // recompute segment sg - 1 into the other set once its last reader is done
            const int bp = (int)((sg - 1) & 1);
            if (rev_issued[bp]) ACHK(hipStreamWaitEvent(s, e_rev[bp], 0));
            if (restore(h, cps[(size_t)(sg - 1)]) != ASPH_OK) return ASPH_ERR;
            do_sort(h, true);
            if (taped_segment(sg - 1) != ASPH_OK) return ASPH_ERR;
        }
        const int b = (int)(sg & 1);
        ACHK(hipStreamWaitEvent(s1, e_rec[b], 0));
        Launch Lr = h->L;   // the segment's work items
        Lr.items = ss[b].items;
        Lr.nitems = ss[b].nitems;
        const int64_t k0 = sg * ck, k1 = std::min<int64_t>(nsteps, k0 + ck);
        for (int64_t k = k1 - 1; k >= k0; --k) {
            const TapeStep& t = tape[b][(size_t)(k - k0)];
            float Ts = 0.f, qs = 0.f;
            sfc_params(Ph, t.simt, Ts, qs);
            launch_adj_step(Lr, S, Pd, (int64_t)N, s1, tape_view(ss[b], t), W, st[cur], st[1 - cur], qc0b, dt, Ts,
                            qs, frac, 1);
            cur = 1 - cur;
        }
        if (sg == 0) {
            launch_adj_step(Lr, S, Pd, (int64_t)N, s1, tape_view(ss[b], tinit), W, st[cur], st[1 - cur], qc0b, dt,
                            0.f, 0.f, frac, 0);
            cur = 1 - cur;
        }
        launch_adj_unperm((int64_t)N, s1, ss[b].perm, st[cur], st[1 - cur]);
        cur = 1 - cur;
        ACHK(hipEventRecord(e_rev[b], s1));
        rev_issued[b] = true;
        ACHK(hipGetLastError());
    }
    ACHK(hipStreamSynchronize(s1));
    ACHK(hipStreamSynchronize(s));
    const double adj_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    // ---- results (particle order of the initial state), then the initial state back -------------
    std::vector<float> tb(4 * N);
    ACHK(hipMemcpy(tb.data(), st[cur].t, N * sizeof(float4), hipMemcpyDeviceToHost));
    if (dJ_dtheta0)
        for (size_t i = 0; i < N; ++i) dJ_dtheta0[i] = tb[4 * i];
    if (dJ_dthq4) std::memcpy(dJ_dthq4, tb.data(), 4 * N * sizeof(float));
    double gq = 0.0;
    ACHK(hipMemcpy(&gq, qc0b, sizeof(double), hipMemcpyDeviceToHost));
    if (restore(h, cps[0]) != ASPH_OK) return ASPH_ERR;
    if (reset_aux(h, true) != ASPH_OK) return ASPH_ERR;
    ACHK(hipStreamSynchronize(s));
    r->J = J;
    r->dJ_dqc0 = gq;
    r->forward_s = fwd_s;
    r->adjoint_s = adj_s;
    r->segments = nseg;
    return ASPH_OK;
}

}  // extern "C"
