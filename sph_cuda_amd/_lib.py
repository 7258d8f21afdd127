"""ctypes binding of the C ABI in include/asph.h (library: sph_cuda_amd/libasph.so).

The structures below mirror the C structs field for field; `load()` checks their sizes against
the library's own sizeof exports before anything else is called.
"""
from __future__ import annotations

import ctypes as C
import os

RH_KNOTS = 8

OK, ERR, CFL_ABORT = 0, 1, 2
NOALLOC_MAX_FLOATS = 1048576   # ASPH_NOALLOC_MAX_FLOATS

F_SURFACE_FLUX = 1 << 0
F_SPONGE = 1 << 1
F_AUTOCONV = 1 << 2
F_ACCRETION = 1 << 3
F_RAIN_EVAP = 1 << 4
F_SEDIMENT = 1 << 5
F_GATHER_ACC = 1 << 6
F_MIN_IMAGE = 1 << 7
F_SORT_EVERY_STEP = 1 << 8

QV_DRY, QV_REF_EXP, QV_RH_TABLE = 0, 1, 2

K_KICK_DRIFT, K_DENSITY_EOS, K_MOMENTUM_KDK, K_RAIN, K_SORT = range(5)
C_COUNT = 16
COUNTER_NAMES = ("dens_one_slab", "dens_slabbed", "dens_stagings", "dens_global", "dens_modeb",
                 "dens_cand", "dens_pairs", "mom_one_slab", "mom_slabbed", "mom_stagings",
                 "mom_global", "mom_modeb", "mom_cand", "mom_pairs", "dens_wave_cand",
                 "mom_wave_cand")
K_COUNT = 5
KERNEL_NAMES = ("kick_drift", "density_eos", "momentum_kdk", "rain", "sort")


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
        ("h", C.c_float), ("stencil", C.c_int32),
        ("th0", C.c_float), ("n2", C.c_float), ("cs", C.c_float),
        ("qv_mode", C.c_int32),
        ("qv0", C.c_float), ("qv_hscale", C.c_float), ("qv_ztop", C.c_float), ("qv_htop", C.c_float),
        ("rh_n", C.c_int32),
        ("rh_z", C.c_float * RH_KNOTS), ("rh_v", C.c_float * RH_KNOTS),
        ("alpha_av", C.c_float), ("h_av", C.c_float),
        ("sponge_sigma", C.c_float), ("sponge_frac", C.c_float), ("dthdt_cap", C.c_float),
        ("c_h", C.c_float), ("c_e", C.c_float), ("beta_moist", C.c_float), ("v_bg", C.c_float),
        ("dts_skin", C.c_float), ("ramp", C.c_float),
        ("qc_auto", C.c_float), ("rate_auto", C.c_float), ("k_accr", C.c_float),
        ("v_rain", C.c_float), ("theta_cap", C.c_float), ("qv_cap", C.c_float),
        ("cfl", C.c_float), ("dt_cap", C.c_float),
        ("max_trip", C.c_int32), ("strict_cfl", C.c_int32),
        ("flags", C.c_uint32),
    ]

    def copy(self) -> "Config":
        c = Config()
        C.memmove(C.byref(c), C.byref(self), C.sizeof(Config))
        return c

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz


class ICSpec(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("reference_init", C.c_int32), ("ref_noise", C.c_float),
        ("jitter", C.c_float), ("bubble_dth", C.c_float), ("bubble_r", C.c_float),
        ("bubble_x", C.c_float), ("bubble_y", C.c_float), ("bubble_z", C.c_float),
        ("noise_sigma", C.c_float), ("noise_len", C.c_float), ("qc_init", C.c_float),
        ("white_noise", C.c_float),
    ]


class RunStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64), ("sorts", C.c_int64), ("trips", C.c_int64), ("dumps", C.c_int64),
        ("simt", C.c_double), ("dt", C.c_float), ("dt_required", C.c_float),
        ("umax", C.c_float * 3), ("wall_s", C.c_double),
    ]


class RunArgs(C.Structure):
    _fields_ = [
        ("sim_end", C.c_double), ("max_steps", C.c_int64), ("dt", C.c_float),
        ("fixed_dt", C.c_int32), ("dump_every", C.c_double),
        ("dump_sx", C.c_int32), ("dump_sy", C.c_int32), ("dump_sz", C.c_int32),
        ("outdir", C.c_char_p), ("validate", C.c_int32), ("verbose", C.c_int32),
    ]


class RunEvent(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("step", C.c_int64), ("sorts", C.c_int64), ("trips", C.c_int64),
        ("dumps", C.c_int64), ("simt", C.c_double), ("wall_s", C.c_double), ("dt", C.c_float),
        ("dt_required", C.c_float), ("umax", C.c_float * 3), ("dump_path", C.c_char_p),
    ]


EV_DUMP, EV_REPORT = 1, 2
RUN_HOOK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.POINTER(RunEvent))


class AdjointArgs(C.Structure):
    _fields_ = [("nsteps", C.c_int64), ("dt", C.c_float), ("checkpoint", C.c_int32)]


class AdjointResult(C.Structure):
    _fields_ = [("J", C.c_double), ("dJ_dqc0", C.c_double), ("forward_s", C.c_double),
                ("adjoint_s", C.c_double), ("segments", C.c_int64)]


class Tuning(C.Structure):
    _fields_ = [("chunk", C.c_int32), ("order", C.c_int32)]


TUNE_MAX = 16


class TuneReport(C.Structure):
    _fields_ = [("n", C.c_int32), ("cand", Tuning * TUNE_MAX), ("ms", C.c_double * TUNE_MAX),
                ("best", C.c_int32), ("items", C.c_uint32 * TUNE_MAX), ("cus", C.c_int32)]


P = C.c_void_p
FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every symbol declared in include/asph.h
SIGNATURES = {
    "asph_sizeof_config": (C.c_size_t, []),
    "asph_sizeof_ic_spec": (C.c_size_t, []),
    "asph_sizeof_run_args": (C.c_size_t, []),
    "asph_sizeof_run_stats": (C.c_size_t, []),
    "asph_config_reference": (C.c_int, [C.POINTER(Config), C.c_int32, C.c_int32, C.c_int32]),
    "asph_config_preset": (C.c_int, [C.c_int32, C.POINTER(Config), C.POINTER(ICSpec), FP, I32P]),
    "asph_config_check": (C.c_int, [C.POINTER(Config)]),
    "asph_ensemble_preset": (C.c_int, [C.c_int32, C.c_uint64, C.POINTER(Config), C.POINTER(ICSpec), FP, I32P]),
    "asph_ic_generate": (C.c_int, [C.POINTER(Config), C.POINTER(ICSpec), FP, FP, FP, FP]),
    "asph_create": (C.c_int, [C.POINTER(Config), P, C.POINTER(P)]),
    "asph_destroy": (C.c_int, [P]),
    "asph_last_error": (C.c_char_p, []),
    "asph_get_config": (C.c_int, [P, C.POINTER(Config)]),
    "asph_num_particles": (C.c_int64, [P]),
    "asph_set_state": (C.c_int, [P, FP, FP, FP, FP]),
    "asph_set_state_device": (C.c_int, [P, P, P, P, P, P]),
    "asph_get_state": (C.c_int, [P, FP, FP, FP, FP, FP, FP, FP]),
    "asph_state_device_ptrs": (C.c_int, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
    "asph_set_acc": (C.c_int, [P, FP]),
    "asph_get_rain": (C.c_int, [P, FP, FP]),
    "asph_set_rain": (C.c_int, [P, FP]),
    "asph_get_rain_fx": (C.c_int, [P, I64P, I64P, I64P]),
    "asph_sort": (C.c_int, [P]),
    "asph_get_sort": (C.c_int, [P, U32P, U32P, U32P, I32P, I32P, I32P]),
    "asph_sort_end_bit": (C.c_int, [P]),
    "asph_force": (C.c_int, [P, C.c_float]),
    "asph_kick_drift": (C.c_int, [P, C.c_float]),
    "asph_density_eos": (C.c_int, [P, C.c_float, C.c_double]),
    "asph_momentum_kdk": (C.c_int, [P, C.c_float]),
    "asph_rain": (C.c_int, [P, C.c_float]),
    "asph_step": (C.c_int, [P, C.c_float, C.c_double, C.c_int32]),
    "asph_max_vel": (C.c_int, [P, FP]),
    "asph_clamp_vel": (C.c_int, [P, C.c_float]),
    "asph_audit": (C.c_int, [P, DP]),
    "asph_run": (C.c_int, [P, C.POINTER(RunArgs), C.POINTER(RunStats)]),
    "asph_run_hooked": (C.c_int, [P, C.POINTER(RunArgs), C.c_int32, RUN_HOOK, C.c_void_p, C.POINTER(RunStats)]),
    "asph_run_device": (C.c_int, [P, C.POINTER(RunArgs), C.c_int32, C.POINTER(RunStats)]),
    "asph_get_clock": (C.c_int, [P, DP, DP, FP]),
    "asph_coarse_dump": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, FP]),
    "asph_coarse_dump_noalloc": (C.c_int, [P, C.c_int32, C.c_int32, C.c_int32, FP]),
    "asph_column_diag": (C.c_int, [P, FP]),
    "asph_write_dump": (C.c_int, [P, C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                  C.c_float, C.c_char_p, C.c_size_t]),
    "asph_ensemble_create": (C.c_int, [C.POINTER(Config), C.c_int32, P, C.POINTER(P)]),
    "asph_member_count": (C.c_int32, [P]),
    "asph_member_range": (C.c_int, [P, C.c_int32, I64P, I64P]),
    "asph_member_max_vel": (C.c_int, [P, C.c_int32, FP]),
    "asph_run_graph": (C.c_int, [P, C.c_int64, C.c_float, C.c_int32, C.POINTER(RunStats)]),
    "asph_set_profiling": (C.c_int, [P, C.c_int32]),
    "asph_kernel_times": (C.c_int, [P, DP, I64P]),
    "asph_synchronize": (C.c_int, [P]),
    "asph_set_counters": (C.c_int, [P, C.c_int32]),
    "asph_adjoint_rain": (C.c_int, [P, C.POINTER(AdjointArgs), FP, FP, C.POINTER(AdjointResult)]),
    "asph_set_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "asph_get_tuning": (C.c_int, [P, C.POINTER(Tuning)]),
    "asph_autotune": (C.c_int, [P, C.c_int32, C.POINTER(TuneReport)]),
    "asph_get_counters": (C.c_int, [P, C.POINTER(C.c_uint64)]),
    "asph_probe_valu": (C.c_int, [C.c_int32, C.c_int32, DP]),
}

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libasph.so")


class AsphError(RuntimeError):
    pass


def load():
    """Loads libasph.so (the gfx950 build); raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise AsphError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for name, cls in (("asph_sizeof_config", Config), ("asph_sizeof_ic_spec", ICSpec),
                      ("asph_sizeof_run_args", RunArgs), ("asph_sizeof_run_stats", RunStats)):
        if getattr(lib, name)() != C.sizeof(cls):
            raise AsphError(f"{cls.__name__} layout does not match the library ({name})")
    _LIB = lib
    return lib


def last_error() -> str:
    msg = load().asph_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> int:
    if rc == OK:
        return rc
    msg = load().asph_last_error()
    msg = msg.decode() if msg else ""
    raise AsphError(f"{what} failed (rc={rc}): {msg}")
