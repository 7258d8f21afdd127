/*
 * asph.h - C ABI of the MI355X-native anisotropic-SPH atmospheric step.
 *
 * The reference (asph_h100.cu) has no library interface: its "step interface" is the set of
 * __global__ kernel signatures driven by the loop body of main() (asph_h100.cu:958-1078), its
 * state layout is the set of device arrays allocated at asph_h100.cu:828-851 and its external
 * contracts are the CLI (asph_h100.cu:756-784) and the dump file (asph_h100.cu:743-747,
 * 1106-1120). Every entry point below names the reference code it replaces.
 *
 * Conventions (all entry points):
 *   - return ASPH_OK (0) on success, ASPH_ERR (1) on an argument or device error,
 *     ASPH_CFL_ABORT (2) when the closed-loop CFL controller aborts
 *     (reference: CK -> exit(1) at asph_h100.cu:41-43, `return 1` at 787, `return 2` at 1065);
 *     the library never calls exit(); asph_last_error() returns the message of the last failure
 *     on the calling thread;
 *   - a handle owns its device memory and is used by one host thread at a time; handles are
 *     independent; all work of a handle is issued in order on the stream given at creation
 *     (NULL = the legacy default stream, as the reference);
 *   - host pointers passed in are borrowed for the duration of the call only;
 *   - particle state uses the reference layout: pos, vel, thq as float4 [x,y,z,w] arrays
 *     (thq = [theta, qv, qc, 0]), rho as float, all of length N = nx*ny*nz, in the current
 *     (sorted) particle order.
 */
#ifndef ASPH_H
#define ASPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASPH_OK 0
#define ASPH_ERR 1
#define ASPH_CFL_ABORT 2

/* Physics / behaviour switches (asph_config.flags). */
#define ASPH_F_SURFACE_FLUX   (1u << 0) /* bulk surface flux, asph_h100.cu:606-626 */
#define ASPH_F_SPONGE         (1u << 1) /* Rayleigh sponge, asph_h100.cu:518-530 */
#define ASPH_F_AUTOCONV       (1u << 2) /* autoconversion onto the rain grid, asph_h100.cu:389-394 */
#define ASPH_F_ACCRETION      (1u << 3) /* Kessler accretion (no reference counterpart) */
#define ASPH_F_RAIN_EVAP      (1u << 4) /* Kessler rain evaporation (no reference counterpart) */
#define ASPH_F_SEDIMENT       (1u << 5) /* rain sedimentation, asph_h100.cu:577-599 */
#define ASPH_F_GATHER_ACC     (1u << 6) /* permute acc with the sort (the reference does not) */
#define ASPH_F_MIN_IMAGE      (1u << 7) /* minimum-image pair separations in x,y (the reference
                                           uses raw differences, asph_h100.cu:315-317, 454) */
#define ASPH_F_SORT_EVERY_STEP (1u << 8) /* re-sort every step instead of displacement gating */

/* Background moisture profile qv_ref(z) used by the ICs and by the buoyancy reference. */
#define ASPH_QV_DRY      0 /* qv_ref = 0 */
#define ASPH_QV_REF_EXP  1 /* qv0*exp(-z/hs)*exp(-max(z-zt,0)/ht), asph_h100.cu:149-153 */
#define ASPH_QV_RH_TABLE 2 /* RH(z) piecewise linear in z times qsat(T_ref(z), p_ref(z)) */

#define ASPH_RH_KNOTS 8

typedef struct asph_config {
    /* lattice / cell grid: cells = particles = nx*ny*nz (asph_h100.cu:790) */
    int32_t nx, ny, nz;
    /* cell sizes [m] = the kernel metric per axis (reference constexpr DX,DY,DZ, :48) */
    float dx, dy, dz;
    /* smoothing length in cell units; support radius 2h cells (reference H_SUP=1, :52) */
    float h;
    /* cells searched either side; 0 selects ceil(2h) (reference STENCIL=int(2*H_SUP), :54) */
    int32_t stencil;
    /* reference sounding (:71-73) */
    float th0, n2, cs;
    /* background moisture (:111, :149-153) */
    int32_t qv_mode;
    float qv0, qv_hscale, qv_ztop, qv_htop;
    int32_t rh_n;
    float rh_z[ASPH_RH_KNOTS], rh_v[ASPH_RH_KNOTS];
    /* dissipation (:84-91) */
    float alpha_av, h_av;
    float sponge_sigma, sponge_frac, dthdt_cap;
    /* surface flux (:100-105) */
    float c_h, c_e, beta_moist, v_bg, dts_skin, ramp;
    /* warm rain (:116-120) plus Kessler accretion rate k2 [1/s] */
    float qc_auto, rate_auto, k_accr, v_rain, theta_cap, qv_cap;
    /* closed-loop CFL (:730-734) */
    float cfl, dt_cap;
    int32_t max_trip, strict_cfl;
    uint32_t flags;
} asph_config;

/* Initial-condition recipe (the reference's kernel_init, asph_h100.cu:210-227, generalised:
 * seeded jitter, bubble, correlated noise, initial cloud). */
typedef struct asph_ic_spec {
    uint64_t seed;
    int32_t reference_init;   /* 1: reproduce kernel_init exactly (cell centres + hash noise) */
    float ref_noise;          /* kernel_init's `noise` amplitude [K] (Config::noise, :741) */
    float jitter;             /* particle jitter U(+-jitter*d) per axis, in cells */
    float bubble_dth;         /* bubble amplitude [K]: dth*cos^2(pi r / 2R) for r < R */
    float bubble_r;           /* R [m] */
    float bubble_x, bubble_y, bubble_z; /* centre [m] */
    float noise_sigma;        /* correlated theta noise std [K] (0 = none) */
    float noise_len;          /* Gaussian correlation length [m] */
    float qc_init;            /* initial cloud water inside the bubble radius [kg/kg] */
    float white_noise;        /* white theta noise std [K] (0 = none) */
} asph_ic_spec;

typedef struct asph_t asph_t;

/* Run-loop statistics (the reference's shutdown summary, asph_h100.cu:1163-1175). */
typedef struct asph_run_stats {
    int64_t steps;
    int64_t sorts;
    int64_t trips;
    int64_t dumps;
    double simt;
    float dt;            /* dt after the last controller update */
    float dt_required;   /* last CFL-required dt */
    float umax[3];       /* last per-axis max |v| */
    double wall_s;       /* host wall time of the loop */
} asph_run_stats;

typedef struct asph_run_args {
    double sim_end;       /* stop when simt >= sim_end (ignored if <= 0) */
    int64_t max_steps;    /* stop after this many steps (ignored if <= 0); asph_run and
                             asph_run_device reject a run with neither bound (a hook of
                             asph_run_hooked may be the only stop instead) */
    float dt;             /* initial dt; <= 0 selects the controller's cap */
    int32_t fixed_dt;     /* 1: keep dt fixed, CFL only monitors and gates the sort */
    double dump_every;    /* sim seconds between dumps (<= 0: none) */
    int32_t dump_sx, dump_sy, dump_sz;
    const char* outdir;   /* dump directory (NULL: no dumps) */
    int32_t validate;     /* FP64 audit at every dump, asph_h100.cu:1126-1144 */
    int32_t verbose;      /* print per-dump lines like asph_h100.cu:1121-1124 */
} asph_run_args;

/* --- configuration ------------------------------------------------------------------- */

/* sizeof(asph_config) etc., for binding-layout checks. */
size_t asph_sizeof_config(void);
size_t asph_sizeof_ic_spec(void);
size_t asph_sizeof_run_args(void);
size_t asph_sizeof_run_stats(void);

/* The reference's compile-time constants (asph_h100.cu:45-120) and Config defaults (:728-742)
 * as a runtime configuration, with the reference's behaviour flags (no acc gather, no
 * minimum image, surface flux and sponge on, autoconversion + sedimentation only). */
int asph_config_reference(asph_config* cfg, int32_t nx, int32_t ny, int32_t nz);

/* Workload presets: 1 = 32^3 dry bubble, 2 = 128^3 moist 2:1, 3 = 256^3 moist 4:1,
 * 4 = 64^3 adjoint window, 5 = ensemble member template (nx,ny,nz overwritten by caller).
 * Fills cfg and (if non-NULL) ic, dt and steps. Presets 1-3 cap the closed-loop controller's dt
 * (dt_cap) at the workload's dt, which the controller lowers when the CFL rule requires. */
int asph_config_preset(int32_t which, asph_config* cfg, asph_ic_spec* ic, float* dt,
                       int32_t* steps);

/* The batched-ensemble workload: n members of the preset-5 template (250 m isotropic lattice,
 * moist, white theta noise 0.1 K), member m with N_m log-uniform in [8192, 131072] from
 * splitmix64(splitmix64(seed) ^ m), nz = 32, nx = ny = round(sqrt(N_m / 32)), and that hash as
 * its noise seed. Fills cfgs[n] and (if non-NULL) ics[n], dt and steps. No reference
 * counterpart: the reference runs one domain per process. */
int asph_ensemble_preset(int32_t n, uint64_t seed, asph_config* cfgs, asph_ic_spec* ics, float* dt,
                         int32_t* steps);

/* Validates cfg (Morton limit 1024 cells per axis, asph_h100.cu:785-788; stencil fits). */
int asph_config_check(const asph_config* cfg);

/* --- initial conditions (host) ------------------------------------------------------------ */

/* Generates the initial state on the host into caller arrays (pos4, vel4, thq4: 4*N floats,
 * rho: N floats), replacing kernel_init (asph_h100.cu:210-227, 891). */
int asph_ic_generate(const asph_config* cfg, const asph_ic_spec* ic, float* pos4, float* vel4,
                     float* thq4, float* rho);

/* --- handle ---------------------------------------------------------------------------- */

/* Allocates the device state (asph_h100.cu:835-862). stream: a hipStream_t or NULL. */
int asph_create(const asph_config* cfg, void* stream, asph_t** out);
int asph_destroy(asph_t* h);
const char* asph_last_error(void);
int asph_get_config(const asph_t* h, asph_config* cfg);
int64_t asph_num_particles(const asph_t* h);

/* Uploads a particle state (host pointers) and clears acc, dthdt, prs, the rain grid, the
 * precipitation field and the sort-gating clock (asph_h100.cu:884-889). */
int asph_set_state(asph_t* h, const float* pos4, const float* vel4, const float* thq4,
                   const float* rho);
/* Same with device pointers (e.g. torch tensors); acc4 may be NULL (then zeroed). */
int asph_set_state_device(asph_t* h, const float* pos4, const float* vel4, const float* thq4,
                          const float* rho, const float* acc4);
/* Downloads state to host arrays; any pointer may be NULL. prs and dthdt are the values of
 * the last force evaluation. */
int asph_get_state(asph_t* h, float* pos4, float* vel4, float* thq4, float* rho, float* acc4,
                   float* prs, float* dthdt);
/* Device pointers of the current state buffers (valid until the next sort or step). Positions
 * written through them count as a new state: run asph_sort before the neighbour stages (their
 * drift bounds cover only the motion of asph_kick_drift since the last sort). */
int asph_state_device_ptrs(asph_t* h, float** pos4, float** vel4, float** thq4, float** rho,
                           float** acc4);
/* Sets acc (host float4 array), e.g. for the first kick after an external state upload. */
int asph_set_acc(asph_t* h, const float* acc4);

/* Rain grid [kg per cell] (reference rainG, asph_h100.cu:841) and accumulated surface
 * precipitation [kg/m^2] (precip2d, :849). Stored as exact 64-bit fixed point internally
 * (units of particle mass * 2^-40); these convert to/from float. */
int asph_get_rain(asph_t* h, float* rain_kg, float* precip_kgm2);
int asph_set_rain(asph_t* h, const float* rain_kg);
/* Raw fixed-point rain state (int64, units pmass*2^-40), for exact tests. */
int asph_get_rain_fx(asph_t* h, int64_t* rain, int64_t* precip, int64_t* clipped);

/* --- the step ---------------------------------------------------------------------------- */

/* Neighbour structure: Morton keys, stable LSD radix sort, gather of pos/vel/thq/rho (and acc
 * with ASPH_F_GATHER_ACC), cell [start,end) table and surface list
 * (asph_h100.cu:962-976; kernels :229-278). */
int asph_sort(asph_t* h);
/* Results of the last sort: keys before sorting, sorted keys, permutation (sorted index ->
 * pre-sort index), cell ranges (2 ints per cell, empty = [0,0)), and the surface list (as a
 * set; its order is unspecified, like the reference's atomic append). Any may be NULL. */
int asph_get_sort(asph_t* h, uint32_t* keys, uint32_t* keys_sorted, uint32_t* perm,
                  int32_t* cell_range, int32_t* surf_list, int32_t* surf_count);
int asph_sort_end_bit(const asph_t* h);

/* Initial force evaluation (asph_h100.cu:913-920): density (lagged Shepard), EOS +
 * microphysics with time step dt, momentum. Writes rho, thq, prs, acc, dthdt. */
int asph_force(asph_t* h, float dt);

/* Stage entry points, each one fused kernel of the step (reference kernels in brackets):
 *   kick_drift   [kernel_kick(dt/2) :535, kernel_drift :545]
 *   density_eos  [kernel_surface_flux :606 (flag), kernel_density :286, kernel_eos_micro :340,
 *                 + Kessler accretion / evaporation (flags)]
 *   momentum_kdk [kernel_momentum :415, kernel_theta_update :562, kernel_kick(dt/2),
 *                 kernel_sponge :518, kernel_max_vel :631]
 *   rain         [kernel_rain_adv :577]
 * simt is the simulated time at the start of the step (surface skin temperature, :989-992). */
int asph_kick_drift(asph_t* h, float dt);
int asph_density_eos(asph_t* h, float dt, double simt);
int asph_momentum_kdk(asph_t* h, float dt);
int asph_rain(asph_t* h, float dt);

/* One KDK step in the reference order (asph_h100.cu:962-1038): an optional sort, then the four
 * stages above. Asynchronous: no host synchronisation. */
int asph_step(asph_t* h, float dt, double simt, int32_t do_sort);

/* Per-axis max |v| after the last momentum_kdk (kernel_max_vel, asph_h100.cu:1043-1046).
 * Synchronises the stream. */
int asph_max_vel(asph_t* h, float out[3]);
/* Rescales |v| <= vmax (kernel_clamp_vel, asph_h100.cu:669-675). */
int asph_clamp_vel(asph_t* h, float vmax);

/* FP64 conservation audit (kernel_audit, asph_h100.cu:650-668): out = [KE, PE, IE, water,
 * pz, mass]; out[6] = rain-grid + precipitation water [kg] (closing the budget the reference
 * audit leaves open). Synchronises. */
int asph_audit(asph_t* h, double out[7]);

/* The reference's main loop (asph_h100.cu:958-1078) with its displacement-gated sort and
 * closed-loop CFL controller (trip counter, velocity clamp, abort -> ASPH_CFL_ABORT), and
 * optional dumps. Continues from the handle's clock (asph_set_state resets it). */
int asph_run(asph_t* h, const asph_run_args* args, asph_run_stats* stats);

/* asph_run with the loop's controller on the device (no reference counterpart; the reference
 * loop of asph_h100.cu:958-1081 without its per-step host synchronisation, :1046): each step's
 * sort decision, CFL rules (dt, trips, velocity clamp, abort) and sort interval run in a
 * one-wave kernel inside a captured two-step graph, and the host reads the loop state every
 * `poll` steps (rounded up to even). Single-domain handles; results and statistics are those of
 * asph_run bit for bit. Dumps and audits are written at the steps asph_run writes them (the device
 * pauses until the host has written them); CFL trip messages are not printed, an abort is
 * (verbose). */
int asph_run_device(asph_t* h, const asph_run_args* args, int32_t poll, asph_run_stats* stats);

/* asph_run with a caller hook, for drivers that print the reference's reports (asph_h100.cu:
 * 1121-1159): the hook runs after every dump time (ASPH_EV_DUMP, with the file written, if any)
 * and every report_every steps (ASPH_EV_REPORT, report_every > 0), with the loop's state. A
 * nonzero return stops the run (ASPH_OK). With a hook, asph_run's own per-dump lines and audit
 * are left to the hook; the CFL messages stay (verbose). */
#define ASPH_EV_DUMP   1
#define ASPH_EV_REPORT 2
typedef struct asph_run_event {
    int32_t kind;
    int64_t step, sorts, trips, dumps;
    double simt;          /* simulated time after the step [s] */
    double wall_s;        /* since the loop started */
    float dt, dt_required;
    float umax[3];        /* max |u|, |v|, |w| after the step */
    const char* dump_path;
} asph_run_event;
typedef int (*asph_run_hook)(void* user, asph_t* h, const asph_run_event* ev);
int asph_run_hooked(asph_t* h, const asph_run_args* args, int32_t report_every, asph_run_hook hook, void* user,
                    asph_run_stats* stats);
/* Simulated time and sort-gating clock of the handle. */
int asph_get_clock(const asph_t* h, double* simt, double* since_sort, float* sort_dt);

/* --- ensembles and graph replay ------------------------------------------------------------ */

/* A handle holding n_members independent domains (no reference counterpart: the reference runs
 * one domain per process). Members are concatenated in member order in every state array
 * (asph_member_range gives each member's slice); all members share the stencil (ceil(2h)) and
 * the ASPH_F_GATHER_ACC choice. Every stage runs as one launch over all members. */
int asph_ensemble_create(const asph_config* cfgs, int32_t n_members, void* stream, asph_t** out);
int32_t asph_member_count(const asph_t* h);
int asph_member_range(const asph_t* h, int32_t m, int64_t* first, int64_t* count);
/* Per-axis max |v| of member m after the last momentum stage (synchronises). */
int asph_member_max_vel(asph_t* h, int32_t m, float out[3]);

/* nsteps KDK steps at a fixed dt replayed from captured hipGraphs (one graph per buffer
 * assignment and sort flag, captured on first use), with no host synchronisation between
 * steps; the simulated time advances on the device. The sort runs every sort_every steps
 * (every step with ASPH_F_SORT_EVERY_STEP; sort_every <= 0 gates it like asph_run, with the
 * interval of the last readback). Single domains and ensembles alike. */
int asph_run_graph(asph_t* h, int64_t nsteps, float dt, int32_t sort_every, asph_run_stats* stats);

/* --- dumps -------------------------------------------------------------------------------- */

/* Coarse block means of [theta,qv,qc,u,v,w] (kernel_coarse_dump, asph_h100.cu:680-703) into
 * host out[cx*cy*cz*6], block index cxx + cx*(cyy + cy*czz); empty blocks are 0. */
int asph_coarse_dump(asph_t* h, int32_t sx, int32_t sy, int32_t sz, float* out);
/* Per-column [mean theta, mean qv, max qc, max |w|] (kernel_column_diag, :704-723). */
int asph_column_diag(asph_t* h, float* out4);
/* Writes "<outdir>/asph_t%06.0fs.bin" in the reference format (asph_h100.cu:1106-1120):
 * 68-byte header <2I10i5f> then cx*cy*cz*6 float32. */
int asph_write_dump(asph_t* h, const char* outdir, int32_t sx, int32_t sy, int32_t sz,
                    double simt, float dt, char* path_out, size_t path_cap);

/* Maximum floats (cx*cy*cz*6) of asph_coarse_dump_noalloc: 4 MiB per GPU. */
#define ASPH_NOALLOC_MAX_FLOATS 1048576
/* asph_coarse_dump without any device allocation (no reference counterpart: for callers that
 * cannot allocate, such as code under HIP graph capture, tight serving loops and harnesses that
 * forbid hipMalloc). Same arguments and results as asph_coarse_dump, bit for bit; the block means
 * are staged in a fixed device array of ASPH_NOALLOC_MAX_FLOATS floats that each GPU holds until
 * the process ends, and a dump larger than that is rejected (ASPH_ERR). Every call on a GPU shares
 * that GPU's array: calls that overlap on one GPU race. The handle's device must be the current one
 * (ASPH_ERR otherwise). Synchronises the handle's stream. */
int asph_coarse_dump_noalloc(asph_t* h, int32_t sx, int32_t sy, int32_t sz, float* out);

/* --- adjoint ------------------------------------------------------------------------------- */

/* Adjoint of the rain mass with respect to the initial state and the autoconversion threshold
 * (no reference counterpart). The forward window starts from the handle's state as set by
 * asph_set_state (sort, force evaluation asph_h100.cu:913-920), runs nsteps KDK steps of the
 * reference loop (asph_h100.cu:962-1038) at fixed dt, and rebuilds the neighbour structure at the
 * start of every segment of `checkpoint` steps (the reference's stale cell lists in between). The
 * objective J is the rain water in the domain at the end of the window [kg] (the sum over columns
 * of the column-integrated rain). The reverse sweep recomputes each segment but the last from its
 * checkpoint (on the handle's stream, overlapped with the reverse of the next segment on a
 * stream of its own) and differentiates every stage: density, EOS + microphysics, momentum,
 * kick/drift, sedimentation. FP32 state, FP64 reductions (rain grid, threshold derivative), the
 * EOS stage's derivative in FP64. Allocates for the duration one checkpoint per segment (52 B
 * per particle), two tapes of checkpoint + 1 steps (about 240 B per particle and step), and the
 * reverse stream. */
typedef struct asph_adjoint_args {
    int64_t nsteps;       /* forward window [steps] */
    float dt;             /* fixed step [s] */
    int32_t checkpoint;   /* steps per segment (0: 10) */
} asph_adjoint_args;

typedef struct asph_adjoint_result {
    double J;             /* rain water in the domain at the end of the window [kg] */
    double dJ_dqc0;       /* dJ / d(autoconversion threshold qc_auto) [kg per kg/kg] */
    double forward_s;     /* wall time of the forward window (checkpoints, taping of the last segment) */
    double adjoint_s;     /* wall time of the reverse sweep (with the recomputation) */
    int64_t segments;
} asph_adjoint_result;

/* Single-domain handles with nx, ny >= 2*stencil+3. dJ_dtheta0 [N] (or NULL) receives dJ/dtheta of
 * the initial state, dJ_dthq4 [4N] (or NULL) dJ/d(theta, qv, qc, -), both in the particle order of
 * the state given to asph_set_state. The handle is left holding that initial state. */
int asph_adjoint_rain(asph_t* h, const asph_adjoint_args* a, float* dJ_dtheta0, float* dJ_dthq4,
                      asph_adjoint_result* r);

/* --- measurement -------------------------------------------------------------------------- */

/* The VALU issue roofline of the current device (no reference counterpart; measurement only):
 * independent non-packed FP32 FMA chains (v_fma_f32 and v_fmac_f32) over 256 wavefronts per CU,
 * `iters` FMAs per chain, the best of `reps` launches of each form after a warm-up. Writes wave64
 * FMA instructions per second (times 128 = FLOP/s). Uses the legacy default stream and allocates
 * for the duration. */
int asph_probe_valu(int32_t iters, int32_t reps, double* wave_fma_per_s);

/* --- tuning of the neighbour kernels ----------------------------------------------------- */

/* Shape of the brick decomposition (no reference counterpart). Each 8x8x8-cell brick's listed
 * particles are split into equal work items of at most `chunk` particles (256..1024), one
 * 512-thread workgroup each; order 1 launches the items heaviest first (estimated cost), order
 * 0 in brick key order. Defaults: chunk 768, order 1. Changing it drops captured graphs. */
typedef struct asph_tuning {
    int32_t chunk;
    int32_t order;
} asph_tuning;

#define ASPH_TUNE_MAX 16
typedef struct asph_tune_report {
    int32_t n;                       /* candidates timed */
    asph_tuning cand[ASPH_TUNE_MAX];
    double ms[ASPH_TUNE_MAX];        /* mean ms of density + momentum on the current state */
    int32_t best;                    /* index of the chosen candidate */
    uint32_t items[ASPH_TUNE_MAX];   /* work items (workgroups) of each candidate on this state */
    int32_t cus;                     /* compute units of the device */
} asph_tune_report;

int asph_set_tuning(asph_t* h, const asph_tuning* t);
int asph_get_tuning(asph_t* h, asph_tuning* t);
/* Times the two neighbour kernels for every candidate shape on the current sorted,
 * force-evaluated state (reps launches each, after one warm-up). It keeps the heaviest-first
 * order unless key order beats all its candidates by more than 10%, and within the order kept the
 * fastest candidate; among the heaviest-first candidates within 3% of their fastest, it keeps
 * instead the largest chunk that gives at least two work items per CU, or the smallest chunk when
 * none does. Reports all. The state
 * (including acc, dthdt and max|v|) is unchanged. Allocates scratch for the duration. */
int asph_autotune(asph_t* h, int32_t reps, asph_tune_report* rep);

/* --- profiling --------------------------------------------------------------------------- */

/* Kernel ids for asph_kernel_times. */
#define ASPH_K_KICK_DRIFT   0
#define ASPH_K_DENSITY_EOS  1
#define ASPH_K_MOMENTUM_KDK 2
#define ASPH_K_RAIN         3
#define ASPH_K_SORT         4
#define ASPH_K_COUNT        5
/* When enabled, every launch of the step kernels is bracketed by HIP events on the handle's
 * stream; asph_kernel_times synchronises and returns per-kernel total ms and launch counts
 * since the last reset. */
int asph_set_profiling(asph_t* h, int32_t enable);
int asph_kernel_times(asph_t* h, double ms[ASPH_K_COUNT], int64_t counts[ASPH_K_COUNT]);
int asph_synchronize(asph_t* h);

/* Work counters of the two neighbour-loop kernels (no reference counterpart; diagnostics of
 * which path each brick took and how much pair work it did). Indices into the array that
 * asph_get_counters fills, summed over launches since the last read: */
#define ASPH_C_DENS_ONE_SLAB  0  /* density: bricks whose halo was staged in LDS at once      */
#define ASPH_C_DENS_SLABBED   1  /*          bricks whose halo was streamed in z-slabs        */
#define ASPH_C_DENS_STAGINGS  2  /*          slab stagings (slabs x rounds of 512 particles) */
#define ASPH_C_DENS_GLOBAL    3  /*          bricks that took the global-memory loop          */
#define ASPH_C_DENS_MODEB     4  /*          work items of the thread-per-particle mode       */
#define ASPH_C_DENS_CAND      5  /*          candidates tested after culling (LDS paths)      */
#define ASPH_C_DENS_PAIRS     6  /*          in-support pairs evaluated (LDS paths)           */
#define ASPH_C_MOM_ONE_SLAB   7  /* momentum: as 0..6                                          */
#define ASPH_C_MOM_SLABBED    8
#define ASPH_C_MOM_STAGINGS   9
#define ASPH_C_MOM_GLOBAL    10
#define ASPH_C_MOM_MODEB     11
#define ASPH_C_MOM_CAND      12
#define ASPH_C_MOM_PAIRS     13
#define ASPH_C_DENS_WAVE_CAND 14 /* density: sum over wavefronts of 64 x the largest per-lane   */
                                 /*          candidate count (lane-slots spent on the walk)     */
#define ASPH_C_MOM_WAVE_CAND 15  /* momentum: as 14                                           */
#define ASPH_C_COUNT         16
/* Enables (allocating and zeroing the counters) or disables counting. */
int asph_set_counters(asph_t* h, int32_t enable);
/* Synchronises, copies the counters to out[ASPH_C_COUNT] and zeroes them. */
int asph_get_counters(asph_t* h, uint64_t out[ASPH_C_COUNT]);

#ifdef __cplusplus
}
#endif

#endif /* ASPH_H */
