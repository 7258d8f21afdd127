// asph_config.cpp - host-only part of the C ABI (include/asph.h): configurations, workload
// presets, validation, initial conditions, and the derived device parameters (make_params).
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

#include "asph_host_common.h"

using namespace asph;

namespace asph {

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

int ceil_log2_ref(int n) { int b = 1; while ((1 << b) < n) ++b; return b; }   // asph_h100.cu:866

int stencil_of(const asph_config& c) {
    return c.stencil > 0 ? c.stencil : (int)std::ceil((double)c.h * 2.0 - 1e-6);
}

}  // namespace asph

namespace {

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
    P.fnx = (float)c.nx; P.fny = (float)c.ny;
    P.hnx = 0.5f * P.fnx; P.hny = 0.5f * P.fny;
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
    P.gmx = P.mm * P.inv_dx; P.gmy = P.mm * P.inv_dy; P.gmz = P.mm * P.inv_dz;
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

extern "C" {

const char* asph_last_error(void) { return asph::g_err.c_str(); }

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
            c->dt_cap = pdt;   // the workload's dt; the CFL controller lowers it when required
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
            c->dt_cap = pdt;
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
            c->dt_cap = pdt;
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

int asph_ensemble_preset(int32_t n, uint64_t seed, asph_config* cfgs, asph_ic_spec* ics, float* dt,
                         int32_t* steps) {
    if (n < 1 || !cfgs) return fail("need n >= 1 and a config array");
    asph_config tmpl;
    asph_ic_spec ic;
    float pdt = 0.f;
    int32_t pst = 0;
    if (asph_config_preset(5, &tmpl, &ic, &pdt, &pst) != ASPH_OK) return ASPH_ERR;
    for (int32_t m = 0; m < n; ++m) {
        // member m: particle count log-uniform in [8192, 131072], lattice nz = 32 and
        // nx = ny = round(sqrt(N / 32)); its own noise seed
        const uint64_t hm = splitmix64(splitmix64(seed) ^ (uint64_t)m);
        const double u = (double)(hm >> 11) * (1.0 / 9007199254740992.0);
        const double N = std::exp(std::log(8192.0) + u * (std::log(131072.0) - std::log(8192.0)));
        const int nxy = (int)std::lround(std::sqrt(N / 32.0));
        cfgs[m] = tmpl;
        cfgs[m].nx = cfgs[m].ny = nxy;
        cfgs[m].nz = 32;
        if (ics) {
            ics[m] = ic;
            ics[m].seed = hm;
        }
    }
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

}  // extern "C"
