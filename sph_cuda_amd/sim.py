"""Python mirror of the reference's step interface over the C ABI (include/asph.h).

`Sim` owns one device handle. Its methods are the reference's step pieces
(asph_h100.cu:958-1078): `sort` (keys, radix sort, gather, cell table), `force` (initial
density/EOS/momentum), the four fused stages, `step`, and `run` (the closed-loop CFL driver).
States are exchanged as numpy arrays in the reference layout: pos, vel, thq of shape (N, 4)
float32 and rho of shape (N,) float32.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


def _fp(a: np.ndarray):
    return a.ctypes.data_as(L.FP)


def preset(which: int):
    """Returns (Config, ICSpec, dt, steps) of a workload preset (1..5)."""
    lib = L.load()
    cfg, ic = L.Config(), L.ICSpec()
    dt, steps = C.c_float(), C.c_int32()
    L.check(lib.asph_config_preset(which, C.byref(cfg), C.byref(ic), C.byref(dt), C.byref(steps)),
            "asph_config_preset")
    return cfg, ic, float(dt.value), int(steps.value)


def ensemble_preset(n: int, seed: int = 5):
    """The batched-ensemble workload: ([Config] * n, [ICSpec] * n, dt, steps)."""
    cfgs, ics = (L.Config * n)(), (L.ICSpec * n)()
    dt, steps = C.c_float(), C.c_int32()
    L.check(L.load().asph_ensemble_preset(n, seed, cfgs, ics, C.byref(dt), C.byref(steps)),
            "asph_ensemble_preset")
    return [c.copy() for c in cfgs], [ics[k] for k in range(n)], float(dt.value), int(steps.value)


def concat_states(states) -> "State":
    """Concatenates member states in member order (the ensemble layout)."""
    return State(*[np.concatenate([getattr(x, f) for x in states]) for f in ("pos", "vel", "thq", "rho")])


def reference_config(nx: int, ny: int, nz: int) -> L.Config:
    lib = L.load()
    cfg = L.Config()
    L.check(lib.asph_config_reference(C.byref(cfg), nx, ny, nz), "asph_config_reference")
    return cfg


def check_config(cfg: L.Config) -> None:
    L.check(L.load().asph_config_check(C.byref(cfg)), "asph_config_check")


@dataclass
class State:
    pos: np.ndarray
    vel: np.ndarray
    thq: np.ndarray
    rho: np.ndarray

    def copy(self) -> "State":
        return State(self.pos.copy(), self.vel.copy(), self.thq.copy(), self.rho.copy())


def ic_generate(cfg: L.Config, ic: L.ICSpec) -> State:
    """Initial state on the host (replaces kernel_init, asph_h100.cu:210-227)."""
    n = cfg.n
    st = State(np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32),
               np.zeros((n, 4), np.float32), np.zeros(n, np.float32))
    L.check(L.load().asph_ic_generate(C.byref(cfg), C.byref(ic), _fp(st.pos), _fp(st.vel),
                                      _fp(st.thq), _fp(st.rho)), "asph_ic_generate")
    return st


class Sim:
    """One device handle: a single domain (cfg) or an ensemble (a list of configs, members
    concatenated in every state array; see member_range)."""

    def __init__(self, cfg, stream: int | None = None):
        self.lib = L.load()
        self.h = C.c_void_p()
        if isinstance(cfg, (list, tuple)):
            arr = (L.Config * len(cfg))(*[c.copy() for c in cfg])
            self.cfgs = [c.copy() for c in cfg]
            L.check(self.lib.asph_ensemble_create(arr, len(cfg), C.c_void_p(stream or 0), C.byref(self.h)),
                    "asph_ensemble_create")
        else:
            self.cfgs = [cfg.copy()]
            L.check(self.lib.asph_create(C.byref(self.cfgs[0]), C.c_void_p(stream or 0), C.byref(self.h)),
                    "asph_create")
        self.cfg = self.cfgs[0]
        self.n = int(self.lib.asph_num_particles(self.h))
        self.ncol = sum(c.nx * c.ny for c in self.cfgs)

    @property
    def members(self) -> int:
        return int(self.lib.asph_member_count(self.h))

    def member_range(self, m: int):
        a, n = C.c_int64(), C.c_int64()
        L.check(self.lib.asph_member_range(self.h, m, C.byref(a), C.byref(n)), "asph_member_range")
        return int(a.value), int(n.value)

    def member_max_vel(self, m: int) -> np.ndarray:
        out = np.empty(3, np.float32)
        L.check(self.lib.asph_member_max_vel(self.h, m, _fp(out)), "asph_member_max_vel")
        return out

    def run_graph(self, nsteps: int, dt: float, sort_every: int = 0):
        """nsteps steps replayed from captured hipGraphs (no host sync between steps)."""
        st = L.RunStats()
        L.check(self.lib.asph_run_graph(self.h, nsteps, dt, sort_every, C.byref(st)), "asph_run_graph")
        return st

    def close(self):
        if self.h:
            self.lib.asph_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state ------------------------------------------------------------------------
    def set_state(self, st: State):
        args = [np.ascontiguousarray(a, np.float32) for a in (st.pos, st.vel, st.thq, st.rho)]
        L.check(self.lib.asph_set_state(self.h, *[_fp(a) for a in args]), "asph_set_state")

    def set_state_device(self, pos, vel, thq, rho, acc=None):
        """Uploads a state from device tensors (torch, contiguous float32 on this GPU): pos, vel and
        thq of shape (N, 4), rho of shape (N,), acc (N, 4) or None (then zeroed)."""
        for name, x, shape in (("pos", pos, (self.n, 4)), ("vel", vel, (self.n, 4)), ("thq", thq, (self.n, 4)),
                               ("rho", rho, (self.n,)), ("acc", acc, (self.n, 4))):
            if x is None:
                continue
            if not (x.is_cuda and x.is_contiguous() and str(x.dtype) == "torch.float32" and tuple(x.shape) == shape):
                raise ValueError(f"{name}: expected a contiguous float32 CUDA tensor of shape {shape}")
        ptr = [None if x is None else C.c_void_p(x.data_ptr()) for x in (pos, vel, thq, rho, acc)]
        L.check(self.lib.asph_set_state_device(self.h, *ptr), "asph_set_state_device")

    def get_state(self) -> State:
        n = self.n
        st = State(np.empty((n, 4), np.float32), np.empty((n, 4), np.float32),
                   np.empty((n, 4), np.float32), np.empty(n, np.float32))
        L.check(self.lib.asph_get_state(self.h, _fp(st.pos), _fp(st.vel), _fp(st.thq), _fp(st.rho),
                                        None, None, None), "asph_get_state")
        return st

    def get_aux(self):
        """(acc (N,4), prs (N,), dthdt (N,)) of the last force evaluation."""
        n = self.n
        acc, prs, dth = np.empty((n, 4), np.float32), np.empty(n, np.float32), np.empty(n, np.float32)
        L.check(self.lib.asph_get_state(self.h, None, None, None, None, _fp(acc), _fp(prs), _fp(dth)),
                "asph_get_state")
        return acc, prs, dth

    def set_acc(self, acc: np.ndarray):
        a = np.ascontiguousarray(acc, np.float32)
        L.check(self.lib.asph_set_acc(self.h, _fp(a)), "asph_set_acc")

    def get_rain_fx(self):
        ncell, ncol = self.n, self.ncol
        r, p, c = np.empty(ncell, np.int64), np.empty(ncol, np.int64), C.c_int64()
        L.check(self.lib.asph_get_rain_fx(self.h, r.ctypes.data_as(L.I64P), p.ctypes.data_as(L.I64P),
                                          C.byref(c)), "asph_get_rain_fx")
        return r, p, int(c.value)

    def get_rain(self):
        ncell, ncol = self.n, self.ncol
        r, p = np.empty(ncell, np.float32), np.empty(ncol, np.float32)
        L.check(self.lib.asph_get_rain(self.h, _fp(r), _fp(p)), "asph_get_rain")
        return r, p

    def set_rain(self, rain_kg: np.ndarray):
        a = np.ascontiguousarray(rain_kg, np.float32)
        L.check(self.lib.asph_set_rain(self.h, _fp(a)), "asph_set_rain")

    # ---- neighbour structure ----------------------------------------------------------------
    def sort(self):
        L.check(self.lib.asph_sort(self.h), "asph_sort")

    def get_sort(self):
        n, ncell = self.n, self.n
        keys, ks, perm = (np.empty(n, np.uint32) for _ in range(3))
        cr = np.empty(2 * ncell, np.int32)
        sl = np.empty(2 * self.ncol, np.int32)
        sc = C.c_int32()
        L.check(self.lib.asph_get_sort(self.h, keys.ctypes.data_as(L.U32P), ks.ctypes.data_as(L.U32P),
                                       perm.ctypes.data_as(L.U32P), cr.ctypes.data_as(L.I32P),
                                       sl.ctypes.data_as(L.I32P), C.byref(sc)), "asph_get_sort")
        return dict(keys=keys, keys_sorted=ks, perm=perm, cell_range=cr.reshape(-1, 2),
                    surf_list=sl[: sc.value].copy())

    # ---- the step -------------------------------------------------------------------------------
    def force(self, dt: float):
        L.check(self.lib.asph_force(self.h, dt), "asph_force")

    def kick_drift(self, dt: float):
        L.check(self.lib.asph_kick_drift(self.h, dt), "asph_kick_drift")

    def density_eos(self, dt: float, simt: float = 0.0):
        L.check(self.lib.asph_density_eos(self.h, dt, simt), "asph_density_eos")

    def momentum_kdk(self, dt: float):
        L.check(self.lib.asph_momentum_kdk(self.h, dt), "asph_momentum_kdk")

    def rain(self, dt: float):
        L.check(self.lib.asph_rain(self.h, dt), "asph_rain")

    def step(self, dt: float, simt: float = 0.0, do_sort: bool = False):
        L.check(self.lib.asph_step(self.h, dt, simt, 1 if do_sort else 0), "asph_step")

    def max_vel(self) -> np.ndarray:
        out = np.empty(3, np.float32)
        L.check(self.lib.asph_max_vel(self.h, _fp(out)), "asph_max_vel")
        return out

    def clamp_vel(self, vmax: float):
        L.check(self.lib.asph_clamp_vel(self.h, vmax), "asph_clamp_vel")

    def audit(self) -> np.ndarray:
        out = np.empty(7, np.float64)
        L.check(self.lib.asph_audit(self.h, out.ctypes.data_as(L.DP)), "asph_audit")
        return out

    def run(self, sim_end: float = 0.0, max_steps: int = 0, dt: float = 0.0, fixed_dt: bool = False,
            dump_every: float = 0.0, dump_stride=(10, 10, 3), outdir: str | None = None,
            validate: bool = False, verbose: bool = False, hook=None, report_every: int = 0):
        """The reference's main loop (asph_h100.cu:958-1078). Returns (rc, RunStats).
        hook(event) (optional) is called with an L.RunEvent after every dump time and every
        report_every steps (asph_run_hooked); a true return value stops the run."""
        a = L.RunArgs()
        a.sim_end, a.max_steps, a.dt, a.fixed_dt = sim_end, max_steps, dt, int(fixed_dt)
        a.dump_every = dump_every
        a.dump_sx, a.dump_sy, a.dump_sz = dump_stride
        a.outdir = outdir.encode() if outdir else None
        a.validate, a.verbose = int(validate), int(verbose)
        st = L.RunStats()
        if hook is None:
            rc = self.lib.asph_run(self.h, C.byref(a), C.byref(st))
        else:
            cb = L.RUN_HOOK(lambda user, h, ev: 1 if hook(ev.contents) else 0)
            rc = self.lib.asph_run_hooked(self.h, C.byref(a), report_every, cb, None, C.byref(st))
        if rc not in (L.OK, L.CFL_ABORT):
            L.check(rc, "asph_run")
        return rc, st

    def run_device(self, sim_end: float = 0.0, max_steps: int = 0, dt: float = 0.0, fixed_dt: bool = False,
                   dump_every: float = 0.0, dump_stride=(10, 10, 3), outdir: str | None = None,
                   validate: bool = False, verbose: bool = False, poll: int = 16):
        """run() with the loop controller on the device (asph_run_device): the same results,
        the host polls every `poll` steps. Returns (rc, RunStats)."""
        a = L.RunArgs()
        a.sim_end, a.max_steps, a.dt, a.fixed_dt = sim_end, max_steps, dt, int(fixed_dt)
        a.dump_every = dump_every
        a.dump_sx, a.dump_sy, a.dump_sz = dump_stride
        a.outdir = outdir.encode() if outdir else None
        a.validate, a.verbose = int(validate), int(verbose)
        st = L.RunStats()
        rc = self.lib.asph_run_device(self.h, C.byref(a), poll, C.byref(st))
        if rc not in (L.OK, L.CFL_ABORT):
            L.check(rc, "asph_run_device")
        return rc, st

    def clock(self):
        t, s, d = C.c_double(), C.c_double(), C.c_float()
        L.check(self.lib.asph_get_clock(self.h, C.byref(t), C.byref(s), C.byref(d)), "asph_get_clock")
        return t.value, s.value, d.value

    # ---- dumps --------------------------------------------------------------------------------
    def coarse_dump(self, sx: int, sy: int, sz: int) -> np.ndarray:
        c = self.cfg
        cx, cy, cz = -(-c.nx // sx), -(-c.ny // sy), -(-c.nz // sz)
        out = np.empty(cx * cy * cz * 6, np.float32)
        L.check(self.lib.asph_coarse_dump(self.h, sx, sy, sz, _fp(out)), "asph_coarse_dump")
        return out.reshape(cz, cy, cx, 6)

    def column_diag(self) -> np.ndarray:
        out = np.empty(self.cfg.nx * self.cfg.ny * 4, np.float32)
        L.check(self.lib.asph_column_diag(self.h, _fp(out)), "asph_column_diag")
        return out.reshape(-1, 4)

    def write_dump(self, outdir: str, stride, simt: float, dt: float) -> str:
        buf = C.create_string_buffer(1024)
        L.check(self.lib.asph_write_dump(self.h, outdir.encode(), stride[0], stride[1], stride[2],
                                         simt, dt, buf, 1024), "asph_write_dump")
        return buf.value.decode()

    # ---- profiling ------------------------------------------------------------------------------
    def set_profiling(self, on: bool):
        L.check(self.lib.asph_set_profiling(self.h, 1 if on else 0), "asph_set_profiling")

    def kernel_times(self):
        ms = np.zeros(L.K_COUNT, np.float64)
        cnt = np.zeros(L.K_COUNT, np.int64)
        L.check(self.lib.asph_kernel_times(self.h, ms.ctypes.data_as(L.DP), cnt.ctypes.data_as(L.I64P)),
                "asph_kernel_times")
        return {L.KERNEL_NAMES[k]: (float(ms[k]), int(cnt[k])) for k in range(L.K_COUNT)}

    def adjoint_rain(self, nsteps: int, dt: float, checkpoint: int = 10) -> dict:
        """Rain water at the end of an nsteps window and its gradient with respect to the initial
        state (as set by set_state) and the autoconversion threshold (include/asph.h,
        asph_adjoint_rain). Returns J [kg], dJ_dqc0, grad_theta (N,), grad_thq (N, 4) and times."""
        a = L.AdjointArgs(nsteps, dt, checkpoint)
        r = L.AdjointResult()
        g = np.empty(self.n, np.float32)
        g4 = np.empty((self.n, 4), np.float32)
        L.check(self.lib.asph_adjoint_rain(self.h, C.byref(a), _fp(g), _fp(g4), C.byref(r)),
                "asph_adjoint_rain")
        return dict(J=r.J, dJ_dqc0=r.dJ_dqc0, grad_theta=g, grad_thq=g4, forward_s=r.forward_s,
                    adjoint_s=r.adjoint_s, segments=r.segments)

    def tuning(self):
        t = L.Tuning()
        L.check(self.lib.asph_get_tuning(self.h, C.byref(t)), "asph_get_tuning")
        return {"chunk": t.chunk, "order": t.order}

    def set_tuning(self, chunk: int, order: int):
        t = L.Tuning(chunk, order)
        L.check(self.lib.asph_set_tuning(self.h, C.byref(t)), "asph_set_tuning")

    def autotune(self, reps: int = 3):
        """Times every brick-decomposition shape on the current state and keeps the fastest.
        Returns {"best": {...}, "trials": [(chunk, order, ms), ...]}."""
        r = L.TuneReport()
        L.check(self.lib.asph_autotune(self.h, reps, C.byref(r)), "asph_autotune")
        trials = [(r.cand[k].chunk, r.cand[k].order, r.ms[k], r.items[k]) for k in range(r.n)]
        return {"best": {"chunk": r.cand[r.best].chunk, "order": r.cand[r.best].order},
                "trials": trials, "cus": r.cus}

    def set_counters(self, on: bool):
        L.check(self.lib.asph_set_counters(self.h, 1 if on else 0), "asph_set_counters")

    def counters(self) -> dict:
        """Work counters of the neighbour kernels since the last call (see include/asph.h)."""
        out = np.zeros(L.C_COUNT, np.uint64)
        L.check(self.lib.asph_get_counters(self.h, out.ctypes.data_as(C.POINTER(C.c_uint64))),
                "asph_get_counters")
        return {n: int(out[k]) for k, n in enumerate(L.COUNTER_NAMES)}

    def synchronize(self):
        L.check(self.lib.asph_synchronize(self.h), "asph_synchronize")
