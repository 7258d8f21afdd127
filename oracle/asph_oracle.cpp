// asph_oracle.cpp - C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// Loaded through ctypes by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
// Each oracle_* function restates the reference code cited beside it; see asph_oracle.hpp.
#include <chrono>
#include <cstdio>
#include <memory>

#include <omp.h>

#include "asph_oracle.hpp"

using namespace oracle;

namespace {

struct Handle {
    int fp64;
    std::unique_ptr<Oracle<float>> f;
    std::unique_ptr<Oracle<double>> d;
};

template <class Fn> int dispatch(Handle* h, Fn&& fn) {
    if (!h) return 1;
    if (h->fp64) fn(*h->d); else fn(*h->f);
    return 0;
}

template <class R> void put4(std::vector<R>& dst, const float* src, int64_t n) {
    for (int64_t i = 0; i < 4 * n; ++i) dst[i] = R(src[i]);
}
template <class R> void get4(const std::vector<R>& src, float* dst, int64_t n) {
    for (int64_t i = 0; i < 4 * n; ++i) dst[i] = (float)src[i];
}

}  // namespace

extern "C" {

void* oracle_create(const asph_config* cfg, int fp64) {
    auto* h = new Handle();
    h->fp64 = fp64;
    if (fp64) h->d.reset(new Oracle<double>(*cfg));
    else h->f.reset(new Oracle<float>(*cfg));
    return h;
}

void oracle_destroy(void* p) { delete (Handle*)p; }

// OpenMP threads the oracle's parallel loops run with (the CPU baseline's core count).
int oracle_threads(void) { return omp_get_max_threads(); }

int oracle_set_state(void* p, const float* pos4, const float* vel4, const float* thq4,
                     const float* rho) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        put4<R>(o.pos, pos4, o.N); put4<R>(o.vel, vel4, o.N); put4<R>(o.thq, thq4, o.N);
        for (int64_t i = 0; i < o.N; ++i) o.rho[i] = R(rho[i]);
        std::fill(o.acc.begin(), o.acc.end(), R(0));
        std::fill(o.prs.begin(), o.prs.end(), R(0));
        std::fill(o.dthdt.begin(), o.dthdt.end(), R(0));
        std::fill(o.rain.begin(), o.rain.end(), 0);
        std::fill(o.dep.begin(), o.dep.end(), 0);
        std::fill(o.precip.begin(), o.precip.end(), 0);
        o.clipped = 0;
        o.simt = 0.0; o.since_sort = 1e9; o.sort_dt = 0.5f; o.initialized = false;
    });
}

// Double-precision state upload (for the FP64 instance's finite differences).
int oracle_set_state_f64(void* p, const double* pos4, const double* vel4, const double* thq4,
                         const double* rho) {
    Handle* h = (Handle*)p;
    if (!h || !h->fp64) return 1;
    auto& o = *h->d;
    for (int64_t i = 0; i < 4 * o.N; ++i) { o.pos[i] = pos4[i]; o.vel[i] = vel4[i]; o.thq[i] = thq4[i]; }
    for (int64_t i = 0; i < o.N; ++i) o.rho[i] = rho[i];
    std::fill(o.acc.begin(), o.acc.end(), 0.0);
    std::fill(o.rain.begin(), o.rain.end(), 0);
    std::fill(o.dep.begin(), o.dep.end(), 0);
    std::fill(o.precip.begin(), o.precip.end(), 0);
    o.simt = 0.0; o.since_sort = 1e9; o.sort_dt = 0.5f; o.initialized = false;
    return 0;
}

int oracle_get_state(void* p, float* pos4, float* vel4, float* thq4, float* rho, float* acc4,
                     float* prs, float* dthdt) {
    return dispatch((Handle*)p, [&](auto& o) {
        if (pos4) get4(o.pos, pos4, o.N);
        if (vel4) get4(o.vel, vel4, o.N);
        if (thq4) get4(o.thq, thq4, o.N);
        if (acc4) get4(o.acc, acc4, o.N);
        for (int64_t i = 0; i < o.N; ++i) {
            if (rho) rho[i] = (float)o.rho[i];
            if (prs) prs[i] = (float)o.prs[i];
            if (dthdt) dthdt[i] = (float)o.dthdt[i];
        }
    });
}

int oracle_get_state_f64(void* p, double* pos4, double* vel4, double* thq4, double* rho) {
    Handle* h = (Handle*)p;
    if (!h || !h->fp64) return 1;
    auto& o = *h->d;
    for (int64_t i = 0; i < 4 * o.N; ++i) {
        if (pos4) pos4[i] = o.pos[i];
        if (vel4) vel4[i] = o.vel[i];
        if (thq4) thq4[i] = o.thq[i];
    }
    if (rho) for (int64_t i = 0; i < o.N; ++i) rho[i] = o.rho[i];
    return 0;
}

int oracle_get_aux_f64(void* p, double* acc4, double* prs, double* dthdt) {
    Handle* h = (Handle*)p;
    if (!h || !h->fp64) return 1;
    auto& o = *h->d;
    for (int64_t i = 0; i < 4 * o.N; ++i) if (acc4) acc4[i] = o.acc[i];
    for (int64_t i = 0; i < o.N; ++i) {
        if (prs) prs[i] = o.prs[i];
        if (dthdt) dthdt[i] = o.dthdt[i];
    }
    return 0;
}

int oracle_set_acc(void* p, const float* acc4) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        put4<R>(o.acc, acc4, o.N);
    });
}

int oracle_set_prs(void* p, const float* prs) {
    return dispatch((Handle*)p, [&](auto& o) {
        for (int64_t i = 0; i < o.N; ++i) o.prs[i] = prs[i];
    });
}

int oracle_set_rain_fx(void* p, const int64_t* rain) {
    return dispatch((Handle*)p, [&](auto& o) {
        for (int64_t c = 0; c < o.ncell; ++c) o.rain[c] = rain[c];
    });
}

int oracle_get_rain_fx(void* p, int64_t* rain, int64_t* precip, int64_t* clipped) {
    return dispatch((Handle*)p, [&](auto& o) {
        if (rain) for (int64_t c = 0; c < o.ncell; ++c) rain[c] = o.rain[c] + o.dep[c];
        if (precip) for (int64_t c = 0; c < o.ncol; ++c) precip[c] = o.precip[c];
        if (clipped) *clipped = o.clipped;
    });
}

int oracle_sort(void* p) { return dispatch((Handle*)p, [&](auto& o) { o.sort(); }); }

int oracle_get_sort(void* p, uint32_t* keys, uint32_t* keys_sorted, uint32_t* perm,
                    int32_t* cell_range, int32_t* surf_list, int32_t* surf_count) {
    return dispatch((Handle*)p, [&](auto& o) {
        for (int64_t i = 0; i < o.N; ++i) {
            if (keys) keys[i] = o.keys[i];
            if (keys_sorted) keys_sorted[i] = o.keys_sorted[i];
            if (perm) perm[i] = o.perm[i];
        }
        if (cell_range) std::copy(o.cell_range.begin(), o.cell_range.end(), cell_range);
        if (surf_list) std::copy(o.surf_list.begin(), o.surf_list.end(), surf_list);
        if (surf_count) *surf_count = (int32_t)o.surf_list.size();
    });
}

int oracle_set_cell_range(void* p, const int32_t* cell_range, const uint32_t* keys_sorted) {
    return dispatch((Handle*)p, [&](auto& o) {
        std::copy(cell_range, cell_range + 2 * o.ncell, o.cell_range.begin());
        if (keys_sorted) for (int64_t i = 0; i < o.N; ++i) o.keys_sorted[i] = keys_sorted[i];
    });
}

int oracle_end_bit(void* p) {
    Handle* h = (Handle*)p;
    return h->fp64 ? h->d->end_bit : h->f->end_bit;
}

// Stage groups that match the device library's fused kernels (reference kernels in order).
int oracle_kick_drift(void* p, float dt) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        std::fill(o.dep.begin(), o.dep.end(), 0);
        o.kick(R(0.5f * dt));
        o.drift(R(dt));
    });
}
int oracle_density_eos(void* p, float dt, double simt) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        if (o.flag(ASPH_F_SURFACE_FLUX)) o.surface_flux(R(dt), simt);
        o.density();
        o.eos_micro(R(dt));
    });
}
int oracle_momentum_kdk(void* p, float dt) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        o.momentum();
        o.theta_update(R(dt));
        o.kick(R(0.5f * dt));
        if (o.flag(ASPH_F_SPONGE)) o.sponge(R(dt));
        o.max_vel(o.umax);
    });
}
int oracle_rain(void* p, float dt) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        o.rain_adv(R(dt));
    });
}
// Individual reference kernels.
int oracle_density(void* p) { return dispatch((Handle*)p, [&](auto& o) { o.density(); }); }
int oracle_eos_micro(void* p, float dt) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        o.eos_micro(R(dt));
    });
}
int oracle_momentum(void* p) { return dispatch((Handle*)p, [&](auto& o) { o.momentum(); }); }
int oracle_force(void* p, float dt) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        o.force(R(dt));
    });
}
int oracle_step(void* p, float dt, double simt, int do_sort) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        o.step(R(dt), simt, do_sort != 0);
    });
}
int oracle_max_vel(void* p, float* out3) {
    return dispatch((Handle*)p, [&](auto& o) { o.max_vel(out3); });
}
int oracle_clamp_vel(void* p, float vmax) {
    return dispatch((Handle*)p, [&](auto& o) {
        using R = typename std::decay_t<decltype(o.pos)>::value_type;
        o.clamp_vel(R(vmax));
    });
}
int oracle_audit(void* p, double* out7) {
    return dispatch((Handle*)p, [&](auto& o) { o.audit(out7); });
}
int oracle_coarse_dump(void* p, int sx, int sy, int sz, float* out) {
    return dispatch((Handle*)p, [&](auto& o) {
        std::vector<float> v;
        o.coarse_dump(sx, sy, sz, v);
        std::copy(v.begin(), v.end(), out);
    });
}
int oracle_column_diag(void* p, float* out) {
    return dispatch((Handle*)p, [&](auto& o) {
        std::vector<float> v;
        o.column_diag(v);
        std::copy(v.begin(), v.end(), out);
    });
}

// The reference main loop with its sort gating and CFL controller (asph_h100.cu:899-921,
// 958-1078), without dumps. Returns 0, or 2 on a CFL abort. stats: [steps, sorts, trips,
// simt, dt, dt_required, ux, uy, uz, wall_s].
int oracle_run(void* p, double sim_end, int64_t max_steps, float dt0, int fixed_dt,
               double* stats) {
    Handle* h = (Handle*)p;
    if (!h || h->fp64) return 1;
    auto& o = *h->f;
    const asph_config& c = o.cfg;
    const float D = std::fmin(std::fmin(c.dx, c.dy), c.dz);
    const float dt_static = c.cfl * D / c.cs;                                   // :794
    const float dt_cap = c.dt_cap > 0.f ? std::fmin(c.dt_cap, dt_static) : dt_static;  // :795
    const float dt_floor = 0.05f * dt_cap;                                      // :796
    float dt = dt0 > 0.f ? dt0 : dt_cap;
    float dt_required = dt;
    if (!o.initialized) {                                                        // :899-921
        o.sort();
        o.force(dt);
    }
    int64_t step = 0, sorts = 0, trips = 0;
    int rc = 0;
    auto t0 = std::chrono::steady_clock::now();
    float u[3] = {0, 0, 0};
    while ((sim_end <= 0 || o.simt < sim_end) && (max_steps <= 0 || step < max_steps)) {
        ++step;
        bool do_sort = (c.flags & ASPH_F_SORT_EVERY_STEP) || o.since_sort >= o.sort_dt;  // :963
        if (do_sort) { o.since_sort = 0.0; ++sorts; }
        o.step(dt, o.simt, do_sort);
        o.simt += dt; o.since_sort += dt;                                        // :1038
        for (int k = 0; k < 3; ++k) u[k] = o.umax[k];
        dt_required = c.cfl * std::fmin(std::fmin(c.dx / (c.cs + u[0]), c.dy / (c.cs + u[1])),
                                         c.dz / (c.cs + u[2])) * 0.90f;         // :1050-1051
        float adopt = fixed_dt ? dt : std::fmin(dt_cap, dt_required);           // :1052
        if (!fixed_dt) {
            if (adopt < dt_floor) {                                              // :1054-1071
                ++trips;
                if (c.strict_cfl || trips > c.max_trip) { rc = 2; break; }
                o.clamp_vel(0.5f * c.cs);
                adopt = std::fmax(dt_floor, c.cfl * D / (1.5f * c.cs) * 0.90f);
            } else if (adopt < 0.98f * dt) {
                trips = 0;                                                       // :1072
            }
            dt = adopt;
        }
        float dz_move = 0.3f * c.dz / std::fmax(u[2], 0.5f);                     // :1075-1077
        float dh_move = 0.3f * c.dx / std::fmax(std::fmax(u[0], u[1]), 0.5f);
        o.sort_dt = std::fmin(std::fmax(std::fmin(dz_move, dh_move), dt), 15.0f);
    }
    double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (stats) {
        stats[0] = (double)step; stats[1] = (double)sorts; stats[2] = (double)trips;
        stats[3] = o.simt; stats[4] = dt; stats[5] = dt_required;
        stats[6] = u[0]; stats[7] = u[1]; stats[8] = u[2]; stats[9] = wall;
    }
    return rc;
}

// Scalar helpers for closed-form tests.
uint32_t oracle_morton3(uint32_t x, uint32_t y, uint32_t z) { return morton3(x, y, z); }
float oracle_fexp(float x) { return asphm::fexp(x); }
float oracle_flog(float x) { return asphm::flog(x); }
float oracle_fpow(float x, float y) { return asphm::fpow(x, y); }

// Scalar physics helpers of a configuration: which = 0 theta_ref, 1 exner_ref, 2 pres_ref,
// 3 rho_ref, 4 qv_ref, 5 esat(T=z), 6 qsat(T=z, p=y), 7 f(q=z) (unnormalised spline),
// 8 kgrad(q=z), 9 dqsdT(T=z, es=esat(z), p=y).
double oracle_scalar(const asph_config* cfg, int fp64, int which, double z, double y) {
    auto eval = [&](auto P, auto zz, auto yy) -> double {
        switch (which) {
            case 0: return (double)P.theta_ref(zz);
            case 1: return (double)P.exner_ref(zz);
            case 2: return (double)P.pres_ref(zz);
            case 3: return (double)P.rho_ref(zz);
            case 4: return (double)P.qv_ref(zz);
            case 5: return (double)P.esat(zz);
            case 6: return (double)P.qsat(zz, yy);
            case 7: return (double)P.fker(zz);
            case 8: return (double)P.kgrad(zz);
            case 9: return (double)P.dqsdT(zz, P.esat(zz), yy);
        }
        return 0.0;
    };
    if (fp64) return eval(Params<double>(*cfg), z, y);
    return eval(Params<float>(*cfg), (float)z, (float)y);
}

}  // extern "C"
