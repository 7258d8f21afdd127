"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg. The product path (sph_cuda_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_LIB = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        build()
    from sph_cuda_amd._lib import Config  # the asph_config layout (include/asph.h)

    lib = C.CDLL(LIB_PATH)
    P, FP, DP = C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_double)
    U32P, I32P, I64P = C.POINTER(C.c_uint32), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    sig = {
        "oracle_create": (P, [C.POINTER(Config), C.c_int]),
        "oracle_destroy": (None, [P]),
        "oracle_threads": (C.c_int, []),
        "oracle_set_state": (C.c_int, [P, FP, FP, FP, FP]),
        "oracle_set_state_f64": (C.c_int, [P, DP, DP, DP, DP]),
        "oracle_get_state": (C.c_int, [P, FP, FP, FP, FP, FP, FP, FP]),
        "oracle_get_state_f64": (C.c_int, [P, DP, DP, DP, DP]),
        "oracle_get_aux_f64": (C.c_int, [P, DP, DP, DP]),
        "oracle_set_acc": (C.c_int, [P, FP]),
        "oracle_set_prs": (C.c_int, [P, FP]),
        "oracle_set_rain_fx": (C.c_int, [P, I64P]),
        "oracle_get_rain_fx": (C.c_int, [P, I64P, I64P, I64P]),
        "oracle_sort": (C.c_int, [P]),
        "oracle_get_sort": (C.c_int, [P, U32P, U32P, U32P, I32P, I32P, I32P]),
        "oracle_set_cell_range": (C.c_int, [P, I32P, U32P]),
        "oracle_end_bit": (C.c_int, [P]),
        "oracle_kick_drift": (C.c_int, [P, C.c_float]),
        "oracle_density_eos": (C.c_int, [P, C.c_float, C.c_double]),
        "oracle_momentum_kdk": (C.c_int, [P, C.c_float]),
        "oracle_rain": (C.c_int, [P, C.c_float]),
        "oracle_density": (C.c_int, [P]),
        "oracle_eos_micro": (C.c_int, [P, C.c_float]),
        "oracle_momentum": (C.c_int, [P]),
        "oracle_force": (C.c_int, [P, C.c_float]),
        "oracle_step": (C.c_int, [P, C.c_float, C.c_double, C.c_int]),
        "oracle_max_vel": (C.c_int, [P, FP]),
        "oracle_clamp_vel": (C.c_int, [P, C.c_float]),
        "oracle_audit": (C.c_int, [P, DP]),
        "oracle_coarse_dump": (C.c_int, [P, C.c_int, C.c_int, C.c_int, FP]),
        "oracle_column_diag": (C.c_int, [P, FP]),
        "oracle_run": (C.c_int, [P, C.c_double, C.c_int64, C.c_float, C.c_int, DP]),
        "oracle_adjoint_rain": (C.c_int, [C.POINTER(Config), DP, DP, DP, DP, C.c_int64, C.c_float, C.c_int,
                                          DP, DP, DP]),
        "oracle_adjoint_rain_f32_trajectory": (C.c_int, [C.POINTER(Config), DP, DP, DP, DP, C.c_int64, C.c_float,
                                                          C.c_int, DP, DP, DP]),
        "oracle_adjoint_set_extra_margin": (None, [C.c_int]),
        "oracle_adjoint_set_float_stages": (None, [C.c_int]),
        "oracle_check_mom_pair_fused": (C.c_double, [C.POINTER(Config), C.c_int, C.c_uint64, C.c_int]),
        "oracle_check_mom_vjp_rev": (C.c_double, [C.POINTER(Config), C.c_int, C.c_uint64, C.c_int, C.c_int]),
        "oracle_morton3": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
        "oracle_fexp": (C.c_float, [C.c_float]),
        "oracle_flog": (C.c_float, [C.c_float]),
        "oracle_fpow": (C.c_float, [C.c_float, C.c_float]),
        "oracle_scalar": (C.c_double, [C.POINTER(Config), C.c_int, C.c_int, C.c_double, C.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


SCALARS = dict(theta_ref=0, exner_ref=1, pres_ref=2, rho_ref=3, qv_ref=4, esat=5, qsat=6,
               fker=7, kgrad=8, dqsdT=9)


def scalar(cfg, name: str, z: float, y: float = 0.0, fp64: bool = False) -> float:
    return load().oracle_scalar(C.byref(cfg), int(fp64), SCALARS[name], z, y)


class Oracle:
    """Same methods as sph_cuda_amd.Sim, computed on the CPU."""

    def __init__(self, cfg, fp64: bool = False):
        self.lib = load()
        self.cfg = cfg.copy()
        self.fp64 = fp64
        self.h = self.lib.oracle_create(C.byref(self.cfg), int(fp64))
        self.n = cfg.nx * cfg.ny * cfg.nz

    def __del__(self):
        try:
            if self.h:
                self.lib.oracle_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_state(self, st):
        a = [np.ascontiguousarray(x, np.float32) for x in (st.pos, st.vel, st.thq, st.rho)]
        self.lib.oracle_set_state(self.h, *[_fp(x) for x in a])

    def set_state_f64(self, pos, vel, thq, rho):
        a = [np.ascontiguousarray(x, np.float64) for x in (pos, vel, thq, rho)]
        D = C.POINTER(C.c_double)
        self.lib.oracle_set_state_f64(self.h, *[x.ctypes.data_as(D) for x in a])

    def get_state(self):
        from sph_cuda_amd.sim import State

        n = self.n
        st = State(np.empty((n, 4), np.float32), np.empty((n, 4), np.float32),
                   np.empty((n, 4), np.float32), np.empty(n, np.float32))
        self.lib.oracle_get_state(self.h, _fp(st.pos), _fp(st.vel), _fp(st.thq), _fp(st.rho),
                                  None, None, None)
        return st

    def get_state_f64(self):
        n = self.n
        D = C.POINTER(C.c_double)
        pos, vel, thq, rho = (np.empty((n, 4)), np.empty((n, 4)), np.empty((n, 4)), np.empty(n))
        self.lib.oracle_get_state_f64(self.h, pos.ctypes.data_as(D), vel.ctypes.data_as(D),
                                      thq.ctypes.data_as(D), rho.ctypes.data_as(D))
        return pos, vel, thq, rho

    def get_aux(self):
        n = self.n
        acc, prs, dth = np.empty((n, 4), np.float32), np.empty(n, np.float32), np.empty(n, np.float32)
        self.lib.oracle_get_state(self.h, None, None, None, None, _fp(acc), _fp(prs), _fp(dth))
        return acc, prs, dth

    def get_aux_f64(self):
        n = self.n
        D = C.POINTER(C.c_double)
        acc, prs, dth = np.empty((n, 4)), np.empty(n), np.empty(n)
        self.lib.oracle_get_aux_f64(self.h, acc.ctypes.data_as(D), prs.ctypes.data_as(D),
                                    dth.ctypes.data_as(D))
        return acc, prs, dth

    def set_acc(self, acc):
        self.lib.oracle_set_acc(self.h, _fp(np.ascontiguousarray(acc, np.float32)))

    def set_prs(self, prs):
        self.lib.oracle_set_prs(self.h, _fp(np.ascontiguousarray(prs, np.float32)))

    def set_rain_fx(self, rain):
        r = np.ascontiguousarray(rain, np.int64)
        self.lib.oracle_set_rain_fx(self.h, r.ctypes.data_as(C.POINTER(C.c_int64)))

    def get_rain_fx(self):
        ncol = self.cfg.nx * self.cfg.ny
        r, p, c = np.empty(self.n, np.int64), np.empty(ncol, np.int64), C.c_int64()
        I = C.POINTER(C.c_int64)
        self.lib.oracle_get_rain_fx(self.h, r.ctypes.data_as(I), p.ctypes.data_as(I), C.byref(c))
        return r, p, int(c.value)

    def sort(self):
        self.lib.oracle_sort(self.h)

    def get_sort(self):
        n = self.n
        U, I = C.POINTER(C.c_uint32), C.POINTER(C.c_int32)
        keys, ks, perm = (np.empty(n, np.uint32) for _ in range(3))
        cr = np.empty(2 * n, np.int32)
        sl = np.empty(2 * self.cfg.nx * self.cfg.ny, np.int32)
        sc = C.c_int32()
        self.lib.oracle_get_sort(self.h, keys.ctypes.data_as(U), ks.ctypes.data_as(U),
                                 perm.ctypes.data_as(U), cr.ctypes.data_as(I), sl.ctypes.data_as(I),
                                 C.byref(sc))
        return dict(keys=keys, keys_sorted=ks, perm=perm, cell_range=cr.reshape(-1, 2),
                    surf_list=sl[: sc.value].copy())

    def set_cell_range(self, cell_range, keys_sorted=None):
        cr = np.ascontiguousarray(cell_range, np.int32).reshape(-1)
        ks = None if keys_sorted is None else np.ascontiguousarray(keys_sorted, np.uint32)
        self.lib.oracle_set_cell_range(self.h, cr.ctypes.data_as(C.POINTER(C.c_int32)),
                                       None if ks is None else ks.ctypes.data_as(C.POINTER(C.c_uint32)))

    def end_bit(self):
        return self.lib.oracle_end_bit(self.h)

    def kick_drift(self, dt):
        self.lib.oracle_kick_drift(self.h, dt)

    def density_eos(self, dt, simt=0.0):
        self.lib.oracle_density_eos(self.h, dt, simt)

    def momentum_kdk(self, dt):
        self.lib.oracle_momentum_kdk(self.h, dt)

    def rain(self, dt):
        self.lib.oracle_rain(self.h, dt)

    def density(self):
        self.lib.oracle_density(self.h)

    def eos_micro(self, dt):
        self.lib.oracle_eos_micro(self.h, dt)

    def momentum(self):
        self.lib.oracle_momentum(self.h)

    def force(self, dt):
        self.lib.oracle_force(self.h, dt)

    def step(self, dt, simt=0.0, do_sort=False):
        self.lib.oracle_step(self.h, dt, simt, int(do_sort))

    def max_vel(self):
        out = np.empty(3, np.float32)
        self.lib.oracle_max_vel(self.h, _fp(out))
        return out

    def clamp_vel(self, vmax):
        self.lib.oracle_clamp_vel(self.h, vmax)

    def audit(self):
        out = np.empty(7, np.float64)
        self.lib.oracle_audit(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def coarse_dump(self, sx, sy, sz):
        c = self.cfg
        cx, cy, cz = -(-c.nx // sx), -(-c.ny // sy), -(-c.nz // sz)
        out = np.empty(cx * cy * cz * 6, np.float32)
        self.lib.oracle_coarse_dump(self.h, sx, sy, sz, _fp(out))
        return out.reshape(cz, cy, cx, 6)

    def column_diag(self):
        out = np.empty(self.cfg.nx * self.cfg.ny * 4, np.float32)
        self.lib.oracle_column_diag(self.h, _fp(out))
        return out.reshape(-1, 4)

    def run(self, sim_end=0.0, max_steps=0, dt=0.0, fixed_dt=False):
        st = np.zeros(10, np.float64)
        rc = self.lib.oracle_run(self.h, sim_end, max_steps, dt, int(fixed_dt),
                                 st.ctypes.data_as(C.POINTER(C.c_double)))
        keys = ("steps", "sorts", "trips", "simt", "dt", "dt_required", "ux", "uy", "uz", "wall_s")
        return rc, dict(zip(keys, st.tolist()))


def adjoint_rain_f64(cfg, st, nsteps, dt, checkpoint=10, extra_margin=0, trajectory="f64", float_stages=0):
    """FP64 host adjoint (oracle/asph_adjoint_ref.cpp) of the rain objective of asph_adjoint_rain,
    same schedule: returns (J [kg], dJ/d(theta, qv, qc, -) of shape (N, 4), dJ/d(qc_auto)).
    extra_margin widens the pair enumeration by that many cells (the result must not change).
    trajectory "f32" runs the reverse (in double) along the FP32 oracle's forward instead.
    float_stages (bits: 1 momentum pairs, 2 density pairs, 4 EOS and microphysics, 8 state adjoints,
    16 momentum epilogue, 32 density ratio)
    evaluates those reverse stages in float, as the device does."""
    lib = load()
    lib.oracle_adjoint_set_extra_margin(int(extra_margin))
    lib.oracle_adjoint_set_float_stages(int(float_stages))
    cfg = cfg.copy()
    arrs = [np.ascontiguousarray(a, np.float64) for a in (st.pos, st.vel, st.thq, st.rho)]
    g = np.zeros((cfg.n, 4), np.float64)
    J, gq = C.c_double(), C.c_double()
    dp = C.POINTER(C.c_double)
    fn = {"f64": lib.oracle_adjoint_rain, "f32": lib.oracle_adjoint_rain_f32_trajectory}[trajectory]
    rc = fn(C.byref(cfg), *[a.ctypes.data_as(dp) for a in arrs], nsteps, dt, checkpoint,
            C.byref(J), g.ctypes.data_as(dp), C.byref(gq))
    lib.oracle_adjoint_set_extra_margin(0)
    lib.oracle_adjoint_set_float_stages(0)
    if rc:
        raise RuntimeError("oracle_adjoint_rain failed")
    return J.value, g, gq.value


def rain_objective_f64(cfg, pos, vel, thq, rho, nsteps, dt, checkpoint=10):
    """The rain objective of asph_adjoint_rain evaluated by the FP64 oracle (same schedule)."""
    o = Oracle(cfg, fp64=True)
    o.set_state_f64(*[np.ascontiguousarray(a, np.float64) for a in (pos, vel, thq, rho)])
    o.sort()
    o.force(dt)
    for k in range(nsteps):
        o.step(dt, k * dt, do_sort=(k > 0 and k % checkpoint == 0))
    r, _, _ = o.get_rain_fx()
    # particle mass as the library evaluates it: rho0 * ((dx * dy) * dz) in float
    f = np.float32
    rho0 = f(f(100000.0) / (f(287.05) * f(cfg.th0)))
    pm = f(rho0 * f(f(f(cfg.dx) * f(cfg.dy)) * f(cfg.dz)))
    return float(np.sum(r.astype(np.float64))) / 2.0**56 * float(pm)   # FP64 rain unit: 2^-56
