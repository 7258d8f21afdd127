// asph_cli.cpp - command-line driver with the reference program's flags, banner, per-dump report,
// audit lines and summary (asph_h100.cu:728-825, 1080-1176), on the C ABI of libasph.
//
// Flags (as the reference's): --nx N --ny N --nz N --dt cap --cfl c --cfl-interval k --max-trip n
// --strict-cfl --dump-every s --sim-hours h --dump-stride sx sz --outdir dir --noise a --validate
// --quick --help. Added: --steps N (stop after N steps), --report N (steps between timing
// reports, the reference's Config::report), --corrected (the workload physics instead of the
// reference's behaviour flags: acc gathered at sorts, minimum image, warm-rain processes on).
// Exit status: 0, 1 on an error, 2 on a CFL abort (asph_h100.cu:1065).
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "asph.h"

namespace {

struct Opts {
    int nx = 1000, ny = 1000, nz = 150;
    float dt_cap = 0.f, cfl = 0.40f;
    int cfl_interval = 8, max_trip = 20;
    bool strict_cfl = false;
    double dump_every = 900.0, sim_hours = 3.0;
    int dump_sx = 10, dump_sy = 10, dump_sz = 3;
    std::string outdir = "dumps";
    int report = 200;
    bool validate = false, corrected = false;
    float noise = 0.1f;
    long long steps = 0;
};

// Per-run state shared with the hook.
struct Report {
    const Opts* o = nullptr;
    asph_config cfg{};
    int ncol = 0;
    double tot0[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<float> coldiag, precip, precip_last;
    double wall_report = 0.0, sim_report = 0.0;
    double simt0 = 0.0;
};

const char* kname(int k) {
    static const char* n[ASPH_K_COUNT] = {"kick+drift", "density+eos", "momentum", "rain", "sort"};
    return n[k];
}

int on_event(void* user, asph_t* h, const asph_run_event* ev) {
    Report& R = *static_cast<Report*>(user);
    if (ev->kind == ASPH_EV_DUMP) {   // asph_h100.cu:1083-1124
        if (asph_column_diag(h, R.coldiag.data()) != ASPH_OK) return 1;
        if (asph_get_rain(h, nullptr, R.precip.data()) != ASPH_OK) return 1;
        int ncloud = 0;
        float qlmax = 0.f, wmax = 0.f;
        double psum = 0.0;
        for (int c = 0; c < R.ncol; ++c) {
            const float ql = R.coldiag[4 * c + 2], w = R.coldiag[4 * c + 3];
            if (ql > 1e-4f) ++ncloud;
            qlmax = std::max(qlmax, ql);
            wmax = std::max(wmax, w);
            psum += (double)R.precip[c] - (double)R.precip_last[c];   // since the last dump
        }
        R.precip_last = R.precip;
        const float cloudcov = R.ncol ? 100.f * (float)ncloud / (float)R.ncol : 0.f;
        printf("%-8lld +%-8.0fs %-8.1f %-9.4f CFL %.2f | cloud %5.2f%% ql_max %.3f g/kg | w_max %5.2f m/s | "
               "precip %.3f mm\n",
               (long long)ev->step, ev->simt, ev->wall_s, ev->dt, ev->dt_required > 0 ? ev->dt / ev->dt_required : 0.f,
               cloudcov, qlmax * 1e3f, wmax, R.ncol ? psum / R.ncol : 0.0);
        if (R.o->validate) {   // asph_h100.cu:1126-1144, with the rain + precipitation water added
            double t[7];
            if (asph_audit(h, t) != ASPH_OK) return 1;
            const double e = t[0] + t[1] + t[2], e0 = R.tot0[0] + R.tot0[1] + R.tot0[2];
            printf("  AUDIT  KE %.4e  PE %.4e  IE %.4e  E_tot %.4e  water %.4e  pz %.4e  rain+precip %.4e\n", t[0],
                   t[1], t[2], e, t[3], t[4], t[6]);
            printf("  DRIFT  dKE %.2e  dE %.2e (%.3f%%/h)  dwater %.2e  dwater_total %.2e\n", t[0] - R.tot0[0],
                   e - e0, 100.0 * (e - e0) / std::fmax(1.0, std::fabs(e0)) * 3600.0 / std::fmax(1.0, ev->simt - R.simt0),
                   t[3] - R.tot0[3], (t[3] + t[6]) - (R.tot0[3] + R.tot0[6]));
        }
        fflush(stdout);
    } else if (ev->kind == ASPH_EV_REPORT) {   // asph_h100.cu:1149-1159
        double ms[ASPH_K_COUNT];
        int64_t cnt[ASPH_K_COUNT];
        if (asph_kernel_times(h, ms, cnt) != ASPH_OK) return 1;
        const double dw = ev->wall_s - R.wall_report, ds = ev->simt - R.sim_report;
        printf("  [step %lld] sim-s/wall-s=%.2f  dt=%.4f  req=%.4f  sorts=%lld  trips=%lld | ", (long long)ev->step,
               dw > 0 ? ds / dw : 0.0, ev->dt, ev->dt_required, (long long)ev->sorts, (long long)ev->trips);
        for (int k = 0; k < ASPH_K_COUNT; ++k)   // ms since the last read (the read resets them)
            if (cnt[k] > 0) printf("%s %.2f  ", kname(k), ms[k]);
        printf("\n");
        fflush(stdout);
        R.wall_report = ev->wall_s;
        R.sim_report = ev->simt;
    }
    return 0;
}

void usage() {
    printf("asph_cli [--nx N --ny N --nz N --dt cap --cfl c\n"
           "          --cfl-interval k --max-trip n --strict-cfl\n"
           "          --dump-every s --sim-hours h --dump-stride sx sz\n"
           "          --outdir dir --noise a --validate --quick\n"
           "          --steps N --report N --corrected]\n");
}

}  // namespace

int main(int argc, char** argv) {
    Opts o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) {
                fprintf(stderr, "ERROR: %s needs a value\n", a.c_str());
                exit(1);
            }
            return argv[++i];
        };
        if (a == "--nx") o.nx = atoi(next());
        else if (a == "--ny") o.ny = atoi(next());
        else if (a == "--nz") o.nz = atoi(next());
        else if (a == "--dt") o.dt_cap = (float)atof(next());
        else if (a == "--cfl") o.cfl = (float)atof(next());
        else if (a == "--cfl-interval") o.cfl_interval = atoi(next());
        else if (a == "--max-trip") o.max_trip = atoi(next());
        else if (a == "--strict-cfl") o.strict_cfl = true;
        else if (a == "--dump-every") o.dump_every = atof(next());
        else if (a == "--sim-hours") o.sim_hours = atof(next());
        else if (a == "--dump-stride") {
            o.dump_sx = o.dump_sy = atoi(next());
            o.dump_sz = atoi(next());
        } else if (a == "--outdir") o.outdir = next();
        else if (a == "--noise") o.noise = (float)atof(next());
        else if (a == "--validate") o.validate = true;
        else if (a == "--quick") {
            o.nx = 128; o.ny = 128; o.nz = 64;
            o.sim_hours = 0.25; o.dump_every = 300.0;
        } else if (a == "--steps") o.steps = atoll(next());
        else if (a == "--report") o.report = atoi(next());
        else if (a == "--corrected") o.corrected = true;
        else if (a == "--help") {
            usage();
            return 0;
        } else {
            fprintf(stderr, "ERROR: unknown flag %s\n", a.c_str());
            usage();
            return 1;
        }
    }
    Report R;
    R.o = &o;
    asph_config& cfg = R.cfg;
    if (asph_config_reference(&cfg, o.nx, o.ny, o.nz) != ASPH_OK) {
        fprintf(stderr, "ERROR: %s\n", asph_last_error());
        return 1;
    }
    cfg.cfl = o.cfl;
    cfg.dt_cap = o.dt_cap;
    cfg.max_trip = o.max_trip;
    cfg.strict_cfl = o.strict_cfl ? 1 : 0;
    if (o.corrected)
        cfg.flags |= ASPH_F_GATHER_ACC | ASPH_F_MIN_IMAGE | ASPH_F_ACCRETION | ASPH_F_RAIN_EVAP;
    if (asph_config_check(&cfg) != ASPH_OK) {   // grid limits, asph_h100.cu:785-788
        fprintf(stderr, "ERROR: %s\n", asph_last_error());
        return 1;
    }
    const int64_t N = (int64_t)o.nx * o.ny * o.nz;
    const float Lx = o.nx * cfg.dx, Ly = o.ny * cfg.dy, Lz = o.nz * cfg.dz;
    const float D = std::fmin(std::fmin(cfg.dx, cfg.dy), cfg.dz);
    const float dt_static = cfg.cfl * D / cfg.cs;
    const float dt_cap = cfg.dt_cap > 0.f ? std::fmin(cfg.dt_cap, dt_static) : dt_static;
    const float dt_floor = 0.05f * dt_cap;

    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        fprintf(stderr, "ERROR: no HIP device\n");
        return 1;
    }
    // particle mass: the reference density P0 / (R_d theta0) times the cell volume (asph_h100.cu:75-76)
    const double pmass = (double)(100000.0f / (287.05f * cfg.th0)) * ((double)cfg.dx * cfg.dy * cfg.dz);
    printf("=== ASPH (gfx950) ==================================================\n");
    printf("device        : %s (%s, %d CU, %.1f GB)\n", prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName,
           prop.multiProcessorCount, prop.totalGlobalMem / 1e9);
    printf("domain        : %.0f x %.0f x %.0f km  periodic xy, reflective z\n", Lx / 1e3, Ly / 1e3, Lz / 1e3);
    printf("parcels       : %d x %d x %d = %lld  (mass %.3e kg each)\n", o.nx, o.ny, o.nz, (long long)N, pmass);
    printf("sounding      : theta0=%.0f K, N^2=%.1e /s^2, c_s=%.0f m/s\n", cfg.th0, cfg.n2, cfg.cs);
    printf("physics       : %s\n", o.corrected ? "corrected (acc gather, minimum image, warm rain)"
                                               : "reference behaviour");
    printf("--- CFL -----------------------------------------------------------\n");
    printf("  static bound : dt <= %.4f s  (CFL %.2f * min(dx,dy,dz)/c_s)\n", dt_static, cfg.cfl);
    printf("  closed loop  : dt <= CFL*Daxis/(c_s+|v|max)*0.9 every %d steps\n", o.cfl_interval);
    printf("  floor %.4f s ; abort after %d violations%s\n", dt_floor, cfg.max_trip, cfg.strict_cfl ? " (STRICT)" : "");
    printf("--- FORCING -------------------------------------------------------\n");
    printf("  bulk flux    : C_H=C_E=%.1e |V|=%.0f m/s beta=%.2f dTs=%.0f K\n", cfg.c_h, cfg.v_bg, cfg.beta_moist,
           cfg.dts_skin);
    printf("  dtheta/dt    : %.2e K/s peak (= %.2f K/h in lowest %.0f m)\n", cfg.c_h * cfg.v_bg * cfg.dts_skin / cfg.dz,
           cfg.c_h * cfg.v_bg * cfg.dts_skin / cfg.dz * 3600.0f, cfg.dz);
    printf("  run          : %.2f sim h, dump every %.0f sim s\n", o.sim_hours, o.dump_every);
    printf("===================================================================\n");
    fflush(stdout);

    asph_t* h = nullptr;
    if (asph_create(&cfg, nullptr, &h) != ASPH_OK) {
        fprintf(stderr, "ERROR: %s\n", asph_last_error());
        return 1;
    }
    std::vector<float> pos(4 * N), vel(4 * N), thq(4 * N), rho(N);
    asph_ic_spec ic;
    std::memset(&ic, 0, sizeof(ic));
    ic.reference_init = 1;   // kernel_init, asph_h100.cu:210-227
    ic.ref_noise = o.noise;
    int rc = asph_ic_generate(&cfg, &ic, pos.data(), vel.data(), thq.data(), rho.data());
    if (rc == ASPH_OK) rc = asph_set_state(h, pos.data(), vel.data(), thq.data(), rho.data());
    if (rc == ASPH_OK) rc = asph_set_profiling(h, 1);
    R.ncol = o.nx * o.ny;
    R.coldiag.assign(4 * (size_t)R.ncol, 0.f);
    R.precip.assign((size_t)R.ncol, 0.f);
    R.precip_last.assign((size_t)R.ncol, 0.f);
    if (rc == ASPH_OK && o.validate) {   // the audit baseline after the first force evaluation
        rc = asph_sort(h);
        if (rc == ASPH_OK) rc = asph_force(h, dt_cap);
        if (rc == ASPH_OK) rc = asph_audit(h, R.tot0);
    }
    if (rc != ASPH_OK) {
        fprintf(stderr, "ERROR: %s\n", asph_last_error());
        asph_destroy(h);
        return 1;
    }
    mkdir(o.outdir.c_str(), 0775);

    asph_run_args a;
    std::memset(&a, 0, sizeof(a));
    a.sim_end = o.sim_hours * 3600.0;
    a.max_steps = o.steps;
    a.dump_every = o.dump_every;
    a.dump_sx = o.dump_sx; a.dump_sy = o.dump_sy; a.dump_sz = o.dump_sz;
    a.outdir = o.outdir.c_str();
    a.validate = o.validate ? 1 : 0;
    a.verbose = 1;
    asph_run_stats st;
    std::memset(&st, 0, sizeof(st));
    rc = asph_run_hooked(h, &a, o.report, on_event, &R, &st);
    if (rc == ASPH_ERR) {
        fprintf(stderr, "ERROR: %s\n", asph_last_error());
        asph_destroy(h);
        return 1;
    }
    const double wall = st.wall_s, rate = wall > 0 ? st.simt / wall : 0.0;   // asph_h100.cu:1163-1175
    printf("\n=== DONE ==========================================================\n");
    printf("wall clock   : %.1f s\n", wall);
    printf("simulated    : %.2f sim h\n", st.simt / 3600.0);
    printf("throughput   : %.2f sim-s per wall-s (%.2fx real time)\n", rate, rate);
    printf("24 h forecast: %.2f wall-h\n", rate > 0 ? 24.0 * 3600.0 / rate / 3600.0 : 0.0);
    printf("parcel-steps : %.2f G/s\n", wall > 0 ? (double)N * (double)st.steps / wall / 1e9 : 0.0);
    printf("sorts %lld, dumps %lld -> %s/\n", (long long)st.sorts, (long long)st.dumps, o.outdir.c_str());
    printf("CFL trips %lld (tolerated %d)\n", (long long)st.trips, cfg.max_trip);
    printf("final dt %.4f s (required %.4f, static %.4f)\n", st.dt, st.dt_required, dt_static);
    printf("===================================================================\n");
    asph_destroy(h);
    return rc == ASPH_CFL_ABORT ? 2 : 0;
}
