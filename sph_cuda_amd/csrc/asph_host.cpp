// asph_host.cpp - the C ABI (include/asph.h): configuration, initial conditions, device state,
// step orchestration, the reference's closed-loop CFL run loop and the dump writer.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "asph_kernels.h"

using namespace asph;

namespace {

thread_local std::string g_err;

int fail(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return ASPH_ERR;
}

#define HIPCHK(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return fail("%s:%d %s: %s", __FILE__, __LINE__, #x,           \
                                          hipGetErrorString(e_));                           \
    } while (0)

#define LAUNCHCHK()                                                                         \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess) return fail("%s:%d launch: %s", __FILE__, __LINE__,           \
                                          hipGetErrorString(e_));                           \
    } while (0)

int ceil_log2_ref(int n) { int b = 1; while ((1 << b) < n) ++b; return b; }   // asph_h100.cu:866

int stencil_of(const asph_config& c) {
    return c.stencil > 0 ? c.stencil : (int)std::ceil((double)c.h * 2.0 - 1e-6);
}

// ---- counter-based RNG (splitmix64) ----------------------------------------------------------
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline double u01(uint64_t seed, uint64_t i, uint64_t stream) {
    uint64_t h = splitmix64(splitmix64(seed) ^ (i * 8u + stream));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
inline double gauss(uint64_t seed, uint64_t i, uint64_t stream) {
    double u1 = u01(seed, i, stream), u2 = u01(seed, i, stream + 1);
    return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
}

// kernel_init's hash noise, asph_h100.cu:187-192.
float hash_noise(int i, int j, int k) {
    uint32_t h = (uint32_t)i * 374761393u + (uint32_t)j * 668265263u + (uint32_t)k * 2246822519u;
    h = (h ^ (h >> 13)) * 1274126177u;
    h ^= h >> 16;
    return (float)(h & 0x00ffffffu) * (1.0f / 16777215.0f) - 0.5f;
}

template <class F> void parallel_for(int64_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 65536) nt = 1;
    std::vector<std::thread> th;
    int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back([=, &f] { for (int64_t i = a; i < b; ++i) f(i); });
    }
    for (auto& x : th) x.join();
}

// Separable Gaussian filter (sigma in cells per axis) of a lattice field; periodic in x, y,
// mirrored in z.
void gauss_filter(std::vector<double>& f, int nx, int ny, int nz, double sx, double sy, double sz) {
    auto kernel = [](double s) {
        int r = std::max(1, (int)std::ceil(3.0 * s));
        std::vector<double> w(2 * r + 1);
        double sum = 0;
        for (int k = -r; k <= r; ++k) { w[k + r] = std::exp(-0.5 * (k / s) * (k / s)); sum += w[k + r]; }
        for (auto& x : w) x /= sum;
        return w;
    };
    std::vector<double> g(f.size());
    if (sx > 0) {
        auto w = kernel(sx);
        int r = ((int)w.size() - 1) / 2;
        parallel_for((int64_t)ny * nz, [&](int64_t row) {
            const double* a = &f[row * nx];
            double* b = &g[row * nx];
            for (int i = 0; i < nx; ++i) {
                double s = 0;
                for (int k = -r; k <= r; ++k) s += w[k + r] * a[((i + k) % nx + nx) % nx];
                b[i] = s;
            }
        });
        f.swap(g);
    }
    if (sy > 0) {
        auto w = kernel(sy);
        int r = ((int)w.size() - 1) / 2;
        parallel_for((int64_t)nz, [&](int64_t iz) {
            for (int iy = 0; iy < ny; ++iy)
                for (int ix = 0; ix < nx; ++ix) {
                    double s = 0;
                    for (int k = -r; k <= r; ++k) {
                        int jy = ((iy + k) % ny + ny) % ny;
                        s += w[k + r] * f[((int64_t)iz * ny + jy) * nx + ix];
                    }
                    g[((int64_t)iz * ny + iy) * nx + ix] = s;
                }
        });
        f.swap(g);
    }
    if (sz > 0) {
        auto w = kernel(sz);
        int r = ((int)w.size() - 1) / 2;
        parallel_for((int64_t)ny, [&](int64_t iy) {
            for (int iz = 0; iz < nz; ++iz)
                for (int ix = 0; ix < nx; ++ix) {
                    double s = 0;
                    for (int k = -r; k <= r; ++k) {
                        int jz = iz + k;
                        // mirror at the walls
                        while (jz < 0 || jz >= nz) jz = jz < 0 ? -jz - 1 : 2 * nz - jz - 1;
                        s += w[k + r] * f[((int64_t)jz * ny + iy) * nx + ix];
                    }
                    g[((int64_t)iz * ny + iy) * nx + ix] = s;
                }
        });
        f.swap(g);
    }
}

}  // namespace

namespace asph {

void make_params(const asph_config& c, DevParams& P) {
    std::memset(&P, 0, sizeof(P));
    P.nx = c.nx; P.ny = c.ny; P.nz = c.nz;
    P.S = stencil_of(c);
    P.N = c.nx * c.ny * c.nz;
    P.ncell = P.N;
    P.ncol = c.nx * c.ny;
    P.end_bit = 3 * std::max(ceil_log2_ref(c.nx), std::max(ceil_log2_ref(c.ny), ceil_log2_ref(c.nz)));
    P.flags = c.flags;
    P.dx = c.dx; P.dy = c.dy; P.dz = c.dz;
    P.inv_dx = 1.0f / c.dx; P.inv_dy = 1.0f / c.dy; P.inv_dz = 1.0f / c.dz;
    P.Lx = (float)c.nx * c.dx; P.Ly = (float)c.ny * c.dy; P.Lz = (float)c.nz * c.dz;
    P.hLx = 0.5f * P.Lx; P.hLy = 0.5f * P.Ly;
    P.inv_h = 1.0f / c.h;
    P.lim = (4.0f * c.h) * c.h;
    float ih5 = (((P.inv_h * P.inv_h) * P.inv_h) * P.inv_h) * P.inv_h;
    P.kin = SIGMA * ih5;
    P.kout = (-0.75f * SIGMA) * ih5;
    P.vcell = (c.dx * c.dy) * c.dz;
    P.th0 = c.th0; P.n2 = c.n2; P.cs = c.cs;
    P.cs2 = c.cs * c.cs;
    P.rho0 = P0 / (R_D * c.th0);
    P.pmass = P.rho0 * P.vcell;
    P.mm = P.pmass / P.vcell;
    P.eta2 = (0.01f * c.h_av) * (0.01f * c.h_av);
    P.visc_c = (-c.alpha_av * c.cs) * c.h_av;
    P.exA = c.n2 > 0.0f ? (GRAV * GRAV) / ((CP * c.th0) * c.n2) : 0.0f;
    P.cp_rd = CP / R_D; P.rd_cp = R_D / CP;
    P.sponge_sigma = c.sponge_sigma; P.sponge_frac = c.sponge_frac;
    P.sponge_zs = (1.0f - c.sponge_frac) * P.Lz;
    P.sponge_depth = c.sponge_frac * P.Lz;
    P.dthdt_cap = c.dthdt_cap;
    P.c_h = c.c_h; P.c_e = c.c_e; P.beta_moist = c.beta_moist; P.v_bg = c.v_bg;
    P.dts_skin = c.dts_skin; P.ramp = c.ramp;
    P.qc_auto = c.qc_auto; P.rate_auto = c.rate_auto; P.k_accr = c.k_accr; P.v_rain = c.v_rain;
    P.theta_cap = c.theta_cap; P.qv_cap = c.qv_cap;
    P.qv_mode = c.qv_mode;
    P.qv0 = c.qv0; P.qv_hs = c.qv_hscale; P.qv_zt = c.qv_ztop; P.qv_ht = c.qv_htop;
    P.rh_n = c.rh_n;
    for (int k = 0; k < ASPH_RH_KNOTS; ++k) { P.rh_z[k] = c.rh_z[k]; P.rh_v[k] = c.rh_v[k]; }
}

}  // namespace asph

struct asph_t {
    asph_config cfg;
    DevParams P;
    hipStream_t stream = nullptr;
    int64_t N = 0, ncell = 0, ncol = 0;
    uint32_t nkeys = 0;
    int nblk = 0;
    float4 *pos = nullptr, *vel = nullptr, *thq = nullptr, *acc = nullptr;
    float4 *pos2 = nullptr, *vel2 = nullptr, *thq2 = nullptr, *acc2 = nullptr;
    float *rho = nullptr, *rho2 = nullptr, *prs = nullptr, *dthdt = nullptr;
    int64_t *rain = nullptr, *rain2 = nullptr, *dep = nullptr, *precip = nullptr;
    unsigned long long* clipped = nullptr;
    uint32_t *keys0 = nullptr, *kA = nullptr, *vA = nullptr, *kB = nullptr, *vB = nullptr;
    uint32_t *bh = nullptr, *bo = nullptr, *dtot = nullptr, *cstart = nullptr;
    uint32_t *keys_sorted = nullptr, *perm = nullptr;
    uint32_t* maxv = nullptr;
    float* disp = nullptr;           // drift since the last sort [cells per axis]
    uint32_t* disp_slots = nullptr;  // per-step drift maxima (64 slots per axis)
    double* audit_buf = nullptr;
    std::vector<void*> allocs;
    bool sorted = false, initialized = false;
    bool dep_pending = false;   // dep holds deposits not yet merged into the rain grid
    double simt = 0.0, since_sort = 1e9;
    float sort_dt = 0.5f;
    // profiling
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_live[ASPH_K_COUNT];

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

namespace {

// Brackets a launch with profiling events when enabled.
struct ProfScope {
    asph_t* h;
    int k;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(asph_t* h_, int k_) : h(h_), k(k_) {
        if (h->prof) {
            a = h->get_event();
            b = h->get_event();
            if (a) hipEventRecord(a, h->stream);
        }
    }
    ~ProfScope() {
        if (h->prof && a && b) {
            hipEventRecord(b, h->stream);
            h->ev_live[k].push_back({a, b});
        }
    }
};

void sfc_params(const DevParams& P, double simt, float& Ts, float& qs) {
    float r = std::fmin(std::fmax((float)simt / P.ramp, 0.f), 1.f);   // asph_h100.cu:989-992
    float sw = (r * r) * (3.f - 2.f * r);
    Ts = P.th0 + P.dts_skin * sw;
    qs = qsat(Ts, pres_ref(P, 0.0f));
}

}  // namespace

extern "C" {

size_t asph_sizeof_config(void) { return sizeof(asph_config); }
size_t asph_sizeof_ic_spec(void) { return sizeof(asph_ic_spec); }
size_t asph_sizeof_run_args(void) { return sizeof(asph_run_args); }
size_t asph_sizeof_run_stats(void) { return sizeof(asph_run_stats); }
const char* asph_last_error(void) { return g_err.c_str(); }

int asph_config_reference(asph_config* c, int32_t nx, int32_t ny, int32_t nz) {
    if (!c) return fail("null config");
    std::memset(c, 0, sizeof(*c));
    c->nx = nx; c->ny = ny; c->nz = nz;
    c->dx = 1000.0f; c->dy = 1000.0f; c->dz = 100.0f;       // :48
    c->h = 1.0f; c->stencil = 2;                             // :52-54
    c->th0 = 300.0f; c->n2 = 6.0e-5f; c->cs = 80.0f;         // :71-73
    c->qv_mode = ASPH_QV_REF_EXP;
    c->qv0 = 0.016f; c->qv_hscale = 2500.0f; c->qv_ztop = 5000.0f; c->qv_htop = 1000.0f;  // :111, 149-153
    c->alpha_av = 0.2f; c->h_av = 2.0f * 100.0f;            // :84-85
    c->sponge_sigma = 0.05f; c->sponge_frac = 0.15f; c->dthdt_cap = 1.0f;   // :89-91
    c->c_h = 1.2e-3f; c->c_e = 1.2e-3f; c->beta_moist = 0.20f; c->v_bg = 7.0f;
    c->dts_skin = 3.15f; c->ramp = 3600.0f;                  // :100-105
    c->qc_auto = 1.0e-3f; c->rate_auto = 1.0e-3f; c->k_accr = 2.2f; c->v_rain = 9.0f;
    c->theta_cap = 500.0f; c->qv_cap = 0.060f;               // :116-120
    c->cfl = 0.40f; c->dt_cap = 0.0f; c->max_trip = 20; c->strict_cfl = 0;  // :730-734
    c->flags = ASPH_F_SURFACE_FLUX | ASPH_F_SPONGE | ASPH_F_AUTOCONV | ASPH_F_SEDIMENT;
    return ASPH_OK;
}

int asph_config_preset(int32_t which, asph_config* c, asph_ic_spec* ic, float* dt, int32_t* steps) {
    if (!c) return fail("null config");
    asph_ic_spec s;
    std::memset(&s, 0, sizeof(s));
    float pdt = 0.f;
    int32_t pst = 0;
    const uint32_t moist = ASPH_F_AUTOCONV | ASPH_F_ACCRETION | ASPH_F_RAIN_EVAP | ASPH_F_SEDIMENT;
    const uint32_t base = ASPH_F_SPONGE | ASPH_F_GATHER_ACC | ASPH_F_MIN_IMAGE;
    auto rh_c2 = [&]() {   // RH 80% below 2 km, linear to 20% at 8 km
        c->qv_mode = ASPH_QV_RH_TABLE;
        c->rh_n = 3;
        c->rh_z[0] = 0.f; c->rh_v[0] = 0.8f;
        c->rh_z[1] = 2000.f; c->rh_v[1] = 0.8f;
        c->rh_z[2] = 8000.f; c->rh_v[2] = 0.2f;
    };
    switch (which) {
        case 1:   // 32^3 dry warm bubble, 8 km cube, isotropic h = 1.3 dx
            asph_config_reference(c, 32, 32, 32);
            c->dx = c->dy = c->dz = 250.f;
            c->h = 1.3f; c->stencil = 0; c->h_av = 2.0f * c->dz;
            c->qv_mode = ASPH_QV_DRY;
            c->flags = base | ASPH_F_AUTOCONV | ASPH_F_SEDIMENT;
            s.seed = 1; s.jitter = 0.1f;
            s.bubble_dth = 2.0f; s.bubble_r = 2000.f;
            s.bubble_x = 4000.f; s.bubble_y = 4000.f; s.bubble_z = 2000.f;
            pdt = 0.5f; pst = 120;
            break;
        case 2:   // 128^3 moist, 16x16x8 km, 2:1 anisotropy
            asph_config_reference(c, 128, 128, 128);
            c->dx = c->dy = 125.f; c->dz = 62.5f;
            c->h = 1.3f; c->stencil = 0; c->h_av = 2.0f * c->dz;
            rh_c2();
            c->flags = base | moist;
            s.seed = 2; s.jitter = 0.1f;
            s.bubble_dth = 1.5f; s.bubble_r = 2000.f;
            s.bubble_x = 8000.f; s.bubble_y = 8000.f; s.bubble_z = 2000.f;
            pdt = 0.25f; pst = 600;
            break;
        case 3:   // 256^3 moist, 32x32x8 km, 4:1 anisotropy, supersaturated 1-2 km layer
            asph_config_reference(c, 256, 256, 256);
            c->dx = c->dy = 125.f; c->dz = 31.25f;
            c->h = 1.3f; c->stencil = 0; c->h_av = 2.0f * c->dz;
            c->qv_mode = ASPH_QV_RH_TABLE;
            c->rh_n = 6;
            c->rh_z[0] = 0.f;    c->rh_v[0] = 0.8f;
            c->rh_z[1] = 1000.f; c->rh_v[1] = 0.8f;
            c->rh_z[2] = 1000.f; c->rh_v[2] = 1.01f;
            c->rh_z[3] = 2000.f; c->rh_v[3] = 1.01f;
            c->rh_z[4] = 2000.f; c->rh_v[4] = 0.8f;
            c->rh_z[5] = 8000.f; c->rh_v[5] = 0.2f;
            c->cfl = 0.6f;   // the 0.2 s step exceeds the 0.4-CFL acoustic bound of 0.156 s
            c->flags = base | moist;
            s.seed = 3; s.jitter = 0.1f;
            s.noise_sigma = 0.2f; s.noise_len = 500.f;
            pdt = 0.2f; pst = 3000;
            break;
        case 4:   // 64^3 adjoint window: 8 km cube, moist, initial cloud in the bubble
            asph_config_reference(c, 64, 64, 64);
            c->dx = c->dy = c->dz = 125.f;
            c->h = 1.3f; c->stencil = 0; c->h_av = 2.0f * c->dz;
            rh_c2();
            c->flags = base | moist | ASPH_F_SORT_EVERY_STEP;
            s.seed = 4; s.jitter = 0.1f;
            s.bubble_dth = 2.0f; s.bubble_r = 2000.f;
            s.bubble_x = 4000.f; s.bubble_y = 4000.f; s.bubble_z = 2000.f;
            s.qc_init = 2.0e-3f;
            pdt = 0.25f; pst = 60;
            break;
        case 5:   // ensemble member template (250 m lattice, nz = 32, moist, white theta noise)
            asph_config_reference(c, 32, 32, 32);
            c->dx = c->dy = c->dz = 250.f;
            c->h = 1.3f; c->stencil = 0; c->h_av = 2.0f * c->dz;
            rh_c2();
            c->flags = base | moist;
            s.seed = 5; s.jitter = 0.1f;
            s.white_noise = 0.1f;
            pdt = 0.25f; pst = 240;
            break;
        default:
            return fail("unknown preset %d", which);
    }
    if (ic) *ic = s;
    if (dt) *dt = pdt;
    if (steps) *steps = pst;
    return ASPH_OK;
}

int asph_config_check(const asph_config* c) {
    if (!c) return fail("null config");
    if (c->nx < 1 || c->ny < 1 || c->nz < 1) return fail("grid must be positive");
    if (c->nx > 1024 || c->ny > 1024 || c->nz > 1024)   // asph_h100.cu:785-788
        return fail("30-bit Morton key addresses 1024 cells/axis");
    int S = stencil_of(*c);
    if (S < 1 || S > 4) return fail("stencil %d outside [1,4]", S);
    if (c->nx < 2 * S + 1 || c->ny < 2 * S + 1)
        return fail("nx, ny must be >= 2*stencil+1 = %d (a wrapped stencil would visit a cell twice)",
                    2 * S + 1);
    if (!(c->dx > 0 && c->dy > 0 && c->dz > 0 && c->h > 0)) return fail("dx, dy, dz, h must be > 0");
    if ((double)2.0 * c->h > S + 1e-6) return fail("stencil %d does not cover the support 2h = %g", S, 2.0 * c->h);
    if (c->qv_mode == ASPH_QV_RH_TABLE && (c->rh_n < 1 || c->rh_n > ASPH_RH_KNOTS))
        return fail("rh_n outside [1,%d]", ASPH_RH_KNOTS);
    for (int k = 1; k < c->rh_n; ++k)
        if (c->rh_z[k] < c->rh_z[k - 1]) return fail("rh_z must be non-decreasing");
    return ASPH_OK;
}

int asph_ic_generate(const asph_config* c, const asph_ic_spec* ic, float* pos4, float* vel4,
                     float* thq4, float* rho) {
    if (asph_config_check(c) != ASPH_OK) return ASPH_ERR;
    if (!ic || !pos4 || !vel4 || !thq4 || !rho) return fail("null argument");
    DevParams P;
    make_params(*c, P);
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const int64_t N = (int64_t)nx * ny * nz;
    std::vector<double> noise;
    if (ic->noise_sigma > 0.f && ic->noise_len > 0.f) {
        noise.resize(N);
        parallel_for(N, [&](int64_t i) { noise[i] = gauss(ic->seed, (uint64_t)i, 3); });
        gauss_filter(noise, nx, ny, nz, ic->noise_len / c->dx, ic->noise_len / c->dy, ic->noise_len / c->dz);
        double s2 = 0;
        for (int64_t i = 0; i < N; ++i) s2 += noise[i] * noise[i];
        double scale = ic->noise_sigma / std::sqrt(std::max(s2 / (double)N, 1e-300));
        for (auto& x : noise) x *= scale;
    }
    parallel_for(N, [&](int64_t i) {
        int ix = (int)(i % nx), iy = (int)((i / nx) % ny), iz = (int)(i / ((int64_t)nx * ny));
        float x, y, z, th, qv, qc = 0.f;
        if (ic->reference_init) {   // kernel_init, asph_h100.cu:217-226
            x = ((float)ix + 0.5f) * c->dx;
            y = ((float)iy + 0.5f) * c->dy;
            z = ((float)iz + 0.5f) * c->dz;
            th = theta_ref(P, z) + ic->ref_noise * hash_noise(ix, iy, iz);
            qv = std::fmax(qv_ref(P, z) * (1.0f + 0.02f * hash_noise(ix + 17, iy + 31, iz + 7)), 0.f);
        } else {
            float jx = (float)((2.0 * u01(ic->seed, (uint64_t)i, 0) - 1.0) * ic->jitter);
            float jy = (float)((2.0 * u01(ic->seed, (uint64_t)i, 1) - 1.0) * ic->jitter);
            float jz = (float)((2.0 * u01(ic->seed, (uint64_t)i, 2) - 1.0) * ic->jitter);
            x = ((float)ix + 0.5f + jx) * c->dx;
            y = ((float)iy + 0.5f + jy) * c->dy;
            z = ((float)iz + 0.5f + jz) * c->dz;
            th = theta_ref(P, z);
            qv = qv_ref(P, z);
            if (ic->bubble_r > 0.f) {
                double ddx = x - ic->bubble_x, ddy = y - ic->bubble_y, ddz = z - ic->bubble_z;
                double r = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                if (r < ic->bubble_r) {
                    double cc = std::cos(0.5 * 3.141592653589793 * r / ic->bubble_r);
                    th += (float)(ic->bubble_dth * cc * cc);
                    qc = ic->qc_init;
                }
            }
            if (!noise.empty()) th += (float)noise[i];
            if (ic->white_noise > 0.f) th += (float)(ic->white_noise * gauss(ic->seed, (uint64_t)i, 5));
        }
        pos4[4 * i] = x; pos4[4 * i + 1] = y; pos4[4 * i + 2] = z; pos4[4 * i + 3] = 0.f;
        vel4[4 * i] = 0.f; vel4[4 * i + 1] = 0.f; vel4[4 * i + 2] = 0.f; vel4[4 * i + 3] = 0.f;
        thq4[4 * i] = th; thq4[4 * i + 1] = qv; thq4[4 * i + 2] = qc; thq4[4 * i + 3] = 0.f;
        rho[i] = rho_ref(P, z);
    });
    return ASPH_OK;
}

int asph_create(const asph_config* cfg, void* stream, asph_t** out) {
    if (!out) return fail("null output");
    *out = nullptr;
    if (asph_config_check(cfg) != ASPH_OK) return ASPH_ERR;
    asph_t* h = new asph_t();
    h->cfg = *cfg;
    make_params(*cfg, h->P);
    h->stream = (hipStream_t)stream;
    h->N = h->P.N; h->ncell = h->P.ncell; h->ncol = h->P.ncol;
    h->nkeys = 1u << h->P.end_bit;
    h->nblk = (int)((h->N + radix_tile() - 1) / radix_tile());
    const size_t N = (size_t)h->N;
    int rc = ASPH_OK;
    rc |= h->alloc(&h->pos, N); rc |= h->alloc(&h->vel, N); rc |= h->alloc(&h->thq, N);
    rc |= h->alloc(&h->acc, N); rc |= h->alloc(&h->pos2, N); rc |= h->alloc(&h->vel2, N);
    rc |= h->alloc(&h->thq2, N); rc |= h->alloc(&h->acc2, N);
    rc |= h->alloc(&h->rho, N); rc |= h->alloc(&h->rho2, N); rc |= h->alloc(&h->prs, N);
    rc |= h->alloc(&h->dthdt, N);
    rc |= h->alloc(&h->rain, (size_t)h->ncell); rc |= h->alloc(&h->rain2, (size_t)h->ncell);
    rc |= h->alloc(&h->dep, (size_t)h->ncell); rc |= h->alloc(&h->precip, (size_t)h->ncol);
    rc |= h->alloc(&h->clipped, 1);
    rc |= h->alloc(&h->keys0, N); rc |= h->alloc(&h->kA, N); rc |= h->alloc(&h->vA, N);
    rc |= h->alloc(&h->kB, N); rc |= h->alloc(&h->vB, N);
    rc |= h->alloc(&h->bh, (size_t)256 * h->nblk); rc |= h->alloc(&h->bo, (size_t)256 * h->nblk);
    rc |= h->alloc(&h->dtot, 256);
    rc |= h->alloc(&h->cstart, (size_t)h->nkeys + 1);
    rc |= h->alloc(&h->maxv, 3 * 64);
    rc |= h->alloc(&h->disp, 4);
    rc |= h->alloc(&h->disp_slots, 3 * 64);
    rc |= h->alloc(&h->audit_buf, 8);
    if (rc != ASPH_OK) { std::string e = g_err; asph_destroy(h); g_err = e; return ASPH_ERR; }
    *out = h;
    return ASPH_OK;
}

int asph_destroy(asph_t* h) {
    if (!h) return ASPH_OK;
    if (h->stream) hipStreamSynchronize(h->stream); else hipDeviceSynchronize();
    for (void* p : h->allocs) hipFree(p);
    for (auto& e : h->ev_pool) hipEventDestroy(e);
    for (auto& v : h->ev_live) for (auto& pr : v) { hipEventDestroy(pr.first); hipEventDestroy(pr.second); }
    delete h;
    return ASPH_OK;
}

int asph_get_config(const asph_t* h, asph_config* cfg) {
    if (!h || !cfg) return fail("null argument");
    *cfg = h->cfg;
    return ASPH_OK;
}

int64_t asph_num_particles(const asph_t* h) { return h ? h->N : 0; }
int asph_sort_end_bit(const asph_t* h) { return h ? h->P.end_bit : 0; }

static int reset_aux(asph_t* h, bool zero_acc) {
    const size_t N = (size_t)h->N;
    if (zero_acc) HIPCHK(hipMemsetAsync(h->acc, 0, N * sizeof(float4), h->stream));
    HIPCHK(hipMemsetAsync(h->prs, 0, N * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->dthdt, 0, N * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->rain, 0, (size_t)h->ncell * sizeof(int64_t), h->stream));
    HIPCHK(hipMemsetAsync(h->dep, 0, (size_t)h->ncell * sizeof(int64_t), h->stream));
    HIPCHK(hipMemsetAsync(h->precip, 0, (size_t)h->ncol * sizeof(int64_t), h->stream));
    HIPCHK(hipMemsetAsync(h->clipped, 0, sizeof(unsigned long long), h->stream));
    HIPCHK(hipMemsetAsync(h->maxv, 0, 3 * 64 * sizeof(uint32_t), h->stream));
    HIPCHK(hipMemsetAsync(h->disp, 0, 4 * sizeof(float), h->stream));
    HIPCHK(hipMemsetAsync(h->disp_slots, 0, 3 * 64 * sizeof(uint32_t), h->stream));
    h->sorted = false;
    h->initialized = false;
    h->dep_pending = false;
    h->simt = 0.0; h->since_sort = 1e9; h->sort_dt = 0.5f;
    return ASPH_OK;
}

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
        r.resize(h->ncell); d.resize(h->ncell);
        HIPCHK(hipMemcpyAsync(r.data(), h->rain, h->ncell * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(d.data(), h->dep, h->ncell * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    }
    if (precip) HIPCHK(hipMemcpyAsync(precip, h->precip, h->ncol * sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
    unsigned long long cl = 0;
    if (clipped) HIPCHK(hipMemcpyAsync(&cl, h->clipped, sizeof(cl), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    // pending deposits (dep) belong to the grid until the next step consumes them
    if (rain) for (int64_t c = 0; c < h->ncell; ++c) rain[c] = r[c] + (h->dep_pending ? d[c] : 0);
    if (clipped) *clipped = (int64_t)cl;
    return ASPH_OK;
}

int asph_get_rain(asph_t* h, float* rain_kg, float* precip_kgm2) {
    if (!h) return fail("null handle");
    std::vector<int64_t> r(rain_kg ? h->ncell : 0), p(precip_kgm2 ? h->ncol : 0);
    if (asph_get_rain_fx(h, rain_kg ? r.data() : nullptr, precip_kgm2 ? p.data() : nullptr, nullptr) != ASPH_OK)
        return ASPH_ERR;
    const double m = (double)h->P.pmass * FX_INV;
    if (rain_kg) for (int64_t c = 0; c < h->ncell; ++c) rain_kg[c] = (float)((double)r[c] * m);
    if (precip_kgm2) {
        const double a = m / ((double)h->P.dx * (double)h->P.dy);
        for (int64_t c = 0; c < h->ncol; ++c) precip_kgm2[c] = (float)((double)p[c] * a);
    }
    return ASPH_OK;
}

int asph_set_rain(asph_t* h, const float* rain_kg) {
    if (!h || !rain_kg) return fail("null argument");
    std::vector<int64_t> r(h->ncell);
    const double inv = FX_SCALE / (double)h->P.pmass;
    for (int64_t c = 0; c < h->ncell; ++c) r[c] = (int64_t)std::llrint((double)rain_kg[c] * inv);
    HIPCHK(hipMemcpyAsync(h->rain, r.data(), h->ncell * sizeof(int64_t), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemsetAsync(h->dep, 0, h->ncell * sizeof(int64_t), h->stream));
    h->dep_pending = false;
    HIPCHK(hipStreamSynchronize(h->stream));
    return ASPH_OK;
}

int asph_sort(asph_t* h) {
    if (!h) return fail("null handle");
    {
        ProfScope ps(h, ASPH_K_SORT);
        launch_radix_sort(h->P, h->stream, h->pos, h->keys0, h->kA, h->vA, h->kB, h->vB, h->bh,
                          h->bo, h->dtot, &h->keys_sorted, &h->perm);
        const bool ga = (h->cfg.flags & ASPH_F_GATHER_ACC) != 0;
        launch_gather(h->stream, (int)h->N, h->perm, h->pos, h->pos2, h->vel, h->vel2, h->thq,
                      h->thq2, h->rho, h->rho2, ga ? h->acc : nullptr, ga ? h->acc2 : nullptr);
        std::swap(h->pos, h->pos2); std::swap(h->vel, h->vel2); std::swap(h->thq, h->thq2);
        std::swap(h->rho, h->rho2);
        if (ga) std::swap(h->acc, h->acc2);
        launch_cell_start(h->stream, (int)h->N, h->nkeys, h->keys_sorted, h->cstart);
        launch_disp_reset(h->stream, h->disp, h->disp_slots);
    }
    LAUNCHCHK();
    h->sorted = true;
    return ASPH_OK;
}

int asph_get_sort(asph_t* h, uint32_t* keys, uint32_t* keys_sorted, uint32_t* perm,
                  int32_t* cell_range, int32_t* surf_list, int32_t* surf_count) {
    if (!h) return fail("null handle");
    if (!h->sorted) return fail("no sort has run on this state");
    const size_t N = (size_t)h->N;
    std::vector<uint32_t> ks(N), cs;
    if (keys) HIPCHK(hipMemcpyAsync(keys, h->keys0, N * 4, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(ks.data(), h->keys_sorted, N * 4, hipMemcpyDeviceToHost, h->stream));
    if (perm) HIPCHK(hipMemcpyAsync(perm, h->perm, N * 4, hipMemcpyDeviceToHost, h->stream));
    if (cell_range) {
        cs.resize((size_t)h->nkeys + 1);
        HIPCHK(hipMemcpyAsync(cs.data(), h->cstart, cs.size() * 4, hipMemcpyDeviceToHost, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    if (keys_sorted) std::memcpy(keys_sorted, ks.data(), N * 4);
    if (cell_range) {   // the reference layout: int2 per row-major cell id, empty = (0,0)
        const DevParams& P = h->P;
        for (int iz = 0; iz < P.nz; ++iz)
            for (int iy = 0; iy < P.ny; ++iy)
                for (int ix = 0; ix < P.nx; ++ix) {
                    int64_t cid = ((int64_t)iz * P.ny + iy) * P.nx + ix;
                    uint32_t k = morton3((uint32_t)ix, (uint32_t)iy, (uint32_t)iz);
                    uint32_t s = cs[k], e = cs[k + 1];
                    cell_range[2 * cid] = s < e ? (int32_t)s : 0;
                    cell_range[2 * cid + 1] = s < e ? (int32_t)e : 0;
                }
    }
    if (surf_list || surf_count) {
        int32_t n = 0;
        int64_t cap = 2 * h->ncol;
        for (size_t i = 0; i < N && n < cap; ++i)
            if ((ks[i] & 0x24924924u) == 0u) { if (surf_list) surf_list[n] = (int32_t)i; ++n; }
        if (surf_count) *surf_count = n;
    }
    return ASPH_OK;
}

static MomArgs mom_args(asph_t* h, float dt, int mode) {
    MomArgs A;
    A.pos = h->pos; A.vel = h->vel; A.thq = h->thq; A.rho = h->rho; A.prs = h->prs;
    A.cstart = h->cstart; A.acc = h->acc; A.dthdt = h->dthdt; A.vel_out = h->vel2;
    A.maxv = h->maxv; A.disp = h->disp; A.disp_slots = h->disp_slots;
    A.dt = dt; A.mode = mode;
    return A;
}

static int density_eos_impl(asph_t* h, float dt, double simt, int bits) {
    if (!h->sorted) return fail("asph_sort must run before the neighbour kernels");
    const DevParams& P = h->P;
    float auto_f = 1.0f - asphm::fexp(-P.rate_auto * dt);
    float Ts = 0.f, qs = 0.f;
    if (bits & 2) sfc_params(P, simt, Ts, qs);
    DensArgs A;
    A.pos = h->pos; A.rho_old = h->rho; A.rho_new = h->rho2; A.thq = h->thq; A.prs = h->prs;
    A.vel = h->vel; A.cstart = h->cstart; A.keys_sorted = h->keys_sorted; A.rain = h->rain;
    A.dep = h->dep; A.maxv = h->maxv; A.disp = h->disp; A.disp_slots = h->disp_slots;
    A.dt = dt; A.auto_f = auto_f; A.Ts = Ts; A.qs_sfc = qs; A.bits = bits;
    {
        ProfScope ps(h, ASPH_K_DENSITY_EOS);
        launch_density_eos(P, h->stream, A);
    }
    LAUNCHCHK();
    std::swap(h->rho, h->rho2);
    if (bits & 1) h->dep_pending = true;
    return ASPH_OK;
}

int asph_force(asph_t* h, float dt) {
    if (!h) return fail("null handle");
    HIPCHK(hipMemsetAsync(h->dep, 0, (size_t)h->ncell * sizeof(int64_t), h->stream));
    if (density_eos_impl(h, dt, 0.0, 1) != ASPH_OK) return ASPH_ERR;
    launch_rain_merge(h->stream, (int)h->ncell, h->rain, h->dep);
    h->dep_pending = false;
    launch_momentum_kdk(h->P, h->stream, mom_args(h, dt, 0));
    LAUNCHCHK();
    h->initialized = true;
    return ASPH_OK;
}

int asph_kick_drift(asph_t* h, float dt) {
    if (!h) return fail("null handle");
    {
        ProfScope ps(h, ASPH_K_KICK_DRIFT);
        launch_kick_drift(h->P, h->stream, h->pos, h->vel, h->acc, h->dep, h->disp, h->disp_slots, dt);
    }
    LAUNCHCHK();
    h->dep_pending = false;
    return ASPH_OK;
}

int asph_density_eos(asph_t* h, float dt, double simt) {
    if (!h) return fail("null handle");
    int bits = 1 | ((h->cfg.flags & ASPH_F_SURFACE_FLUX) ? 2 : 0);
    return density_eos_impl(h, dt, simt, bits);
}

int asph_momentum_kdk(asph_t* h, float dt) {
    if (!h) return fail("null handle");
    if (!h->sorted) return fail("asph_sort must run before the neighbour kernels");
    {
        ProfScope ps(h, ASPH_K_MOMENTUM_KDK);
        launch_momentum_kdk(h->P, h->stream, mom_args(h, dt, 1));
    }
    LAUNCHCHK();
    std::swap(h->vel, h->vel2);
    return ASPH_OK;
}

int asph_rain(asph_t* h, float dt) {
    if (!h) return fail("null handle");
    const DevParams& P = h->P;
    float frac = (P.v_rain * dt) * P.inv_dz;   // asph_h100.cu:587-588
    if (frac > 1.f) frac = 1.f;
    {
        ProfScope ps(h, ASPH_K_RAIN);
        launch_rain(P, h->stream, h->rain, h->dep, h->rain2, h->precip, h->clipped, frac,
                    (P.flags & ASPH_F_SEDIMENT) ? 1 : 0);
    }
    LAUNCHCHK();
    std::swap(h->rain, h->rain2);
    h->dep_pending = false;
    return ASPH_OK;
}

int asph_step(asph_t* h, float dt, double simt, int32_t do_sort) {
    if (!h) return fail("null handle");
    if (do_sort && asph_sort(h) != ASPH_OK) return ASPH_ERR;
    if (asph_kick_drift(h, dt) != ASPH_OK) return ASPH_ERR;
    if (asph_density_eos(h, dt, simt) != ASPH_OK) return ASPH_ERR;
    if (asph_momentum_kdk(h, dt) != ASPH_OK) return ASPH_ERR;
    if (asph_rain(h, dt) != ASPH_OK) return ASPH_ERR;
    return ASPH_OK;
}

int asph_max_vel(asph_t* h, float out[3]) {
    if (!h || !out) return fail("null argument");
    uint32_t m[192];
    HIPCHK(hipMemcpyAsync(m, h->maxv, sizeof(m), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int a = 0; a < 3; ++a) {
        uint32_t mx = 0;
        for (int s = 0; s < 64; ++s) mx = std::max(mx, m[64 * a + s]);
        float f;
        std::memcpy(&f, &mx, 4);
        out[a] = f;
    }
    return ASPH_OK;
}

int asph_clamp_vel(asph_t* h, float vmax) {
    if (!h) return fail("null handle");
    launch_clamp_vel(h->stream, (int)h->N, h->vel, vmax);
    LAUNCHCHK();
    return ASPH_OK;
}

int asph_audit(asph_t* h, double out[7]) {
    if (!h || !out) return fail("null argument");
    HIPCHK(hipMemsetAsync(h->audit_buf, 0, 8 * sizeof(double), h->stream));
    launch_audit(h->P, h->stream, h->pos, h->vel, h->thq, h->rain, h->dep_pending ? h->dep : nullptr,
                 h->precip, h->audit_buf);
    LAUNCHCHK();
    double t[8];
    HIPCHK(hipMemcpyAsync(t, h->audit_buf, sizeof(t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double m = (double)h->P.pmass;
    for (int k = 0; k < 5; ++k) out[k] = t[k] * m;
    out[5] = m * (double)h->N;
    out[6] = t[5] * FX_INV * m;
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
    const DevParams& P = h->P;
    int cx = (P.nx + sx - 1) / sx, cy = (P.ny + sy - 1) / sy, cz = (P.nz + sz - 1) / sz;
    size_t n = (size_t)cx * cy * cz * 6;
    float* d = nullptr;
    HIPCHK(hipMalloc(&d, n * sizeof(float)));
    launch_coarse_dump(P, h->stream, h->cstart, h->vel, h->thq, d, sx, sy, sz, cx, cy, cz);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, n * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d);
    if (e != hipSuccess) return fail("coarse dump: %s", hipGetErrorString(e));
    return ASPH_OK;
}

int asph_column_diag(asph_t* h, float* out4) {
    if (!h || !out4) return fail("null argument");
    if (!h->sorted) return fail("no cell table: run asph_sort first");
    float4* d = nullptr;
    HIPCHK(hipMalloc(&d, (size_t)h->ncol * sizeof(float4)));
    launch_column_diag(h->P, h->stream, h->cstart, h->thq, h->vel, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out4, d, (size_t)h->ncol * sizeof(float4), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d);
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
    const DevParams& P = h->P;
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

int asph_run(asph_t* h, const asph_run_args* a, asph_run_stats* st) {
    if (!h || !a) return fail("null argument");
    const asph_config& c = h->cfg;
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
    double tot0[7] = {0};
    if (a->validate && asph_audit(h, tot0) != ASPH_OK) return ASPH_ERR;
    int64_t step = 0, sorts = 0, trips = 0, dumps = 0;
    float u[3] = {0, 0, 0};
    int rc = ASPH_OK;
    auto t0 = std::chrono::steady_clock::now();
    while ((a->sim_end <= 0 || h->simt < a->sim_end) && (a->max_steps <= 0 || step < a->max_steps)) {
        ++step;
        bool do_sort = (c.flags & ASPH_F_SORT_EVERY_STEP) || h->since_sort >= h->sort_dt;   // :963
        if (do_sort) { h->since_sort = 0.0; ++sorts; }
        if (asph_step(h, dt, h->simt, do_sort ? 1 : 0) != ASPH_OK) return ASPH_ERR;
        h->simt += dt; h->since_sort += dt;                                             // :1038
        if (asph_max_vel(h, u) != ASPH_OK) return ASPH_ERR;                             // :1043-1046
        dt_required = c.cfl * std::fmin(std::fmin(c.dx / (c.cs + u[0]), c.dy / (c.cs + u[1])),
                                        c.dz / (c.cs + u[2])) * 0.90f;
        float adopt = a->fixed_dt ? dt : std::fmin(dt_cap, dt_required);
        if (!a->fixed_dt) {
            if (adopt < dt_floor) {                                                     // :1054-1071
                ++trips;
                if (a->verbose)
                    fprintf(stderr, "[CFL] step %lld t=%.1fs : required %.4f s < floor %.4f s\n",
                            (long long)step, h->simt, adopt, dt_floor);
                if (c.strict_cfl || trips > c.max_trip) {
                    if (a->verbose) fprintf(stderr, "[CFL] ABORT after %lld violation(s).\n", (long long)trips);
                    rc = ASPH_CFL_ABORT;
                    break;
                }
                if (asph_clamp_vel(h, 0.5f * c.cs) != ASPH_OK) return ASPH_ERR;
                adopt = std::fmax(dt_floor, c.cfl * D / (1.5f * c.cs) * 0.90f);
            } else if (adopt < 0.98f * dt) {
                trips = 0;                                                              // :1072
            }
            dt = adopt;
        }
        float dz_move = 0.3f * c.dz / std::fmax(u[2], 0.5f);                            // :1075-1077
        float dh_move = 0.3f * c.dx / std::fmax(std::fmax(u[0], u[1]), 0.5f);
        h->sort_dt = std::fmin(std::fmax(std::fmin(dz_move, dh_move), dt), 15.0f);
        if (h->simt >= next_dump) {                                                     // :1081-1147
            if (a->outdir) {
                char path[1024];
                if (asph_write_dump(h, a->outdir, a->dump_sx, a->dump_sy, a->dump_sz, h->simt, dt,
                                    path, sizeof(path)) != ASPH_OK)
                    return ASPH_ERR;
                ++dumps;
                if (a->verbose) printf("%-8lld +%-8.0fs dt %.4f  CFL %.2f  -> %s\n", (long long)step,
                                       h->simt, dt, dt_required > 0 ? dt / dt_required : 0.f, path);
            }
            if (a->validate) {
                double tt[7];
                if (asph_audit(h, tt) != ASPH_OK) return ASPH_ERR;
                if (a->verbose)
                    printf("  AUDIT  KE %.4e  PE %.4e  IE %.4e  E_tot %.4e  water %.4e  rain %.4e  pz %.4e\n",
                           tt[0], tt[1], tt[2], tt[0] + tt[1] + tt[2], tt[3], tt[6], tt[4]);
            }
            next_dump += a->dump_every;
        }
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

int asph_synchronize(asph_t* h) {
    if (!h) return fail("null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return ASPH_OK;
}

}  // extern "C"
