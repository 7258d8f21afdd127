#!/usr/bin/env python3
"""Throughput of the anisotropic-SPH step on one MI355X.

Workload: the 256^3 = 16,777,216-particle moist step (32 x 32 x 8 km, 4:1 anisotropic kernel
metric, supersaturated 1-2 km layer, Gaussian-correlated theta noise, dt = 0.2 s), synthetic
initial state from the seeded generator, resident in HBM before timing starts. A "step" is one
KDK step of the reference loop (asph_h100.cu:958-1078): displacement-gated sort, kick + drift,
density + EOS + warm-rain microphysics, momentum + theta update + kick + sponge, rain
sedimentation, and the per-step CFL/sort-gating readback. dt is held fixed at 0.2 s (the CFL
controller monitors only).

Prints one JSON line: particle-steps/s, the HBM-roofline fraction of the dominant kernel from
HIP events on the library's stream, the CPU oracle on a bounded sample of the same workload, the
batched ensemble against sequential runs, and the adjoint/forward time ratio of the 64^3 rain
adjoint.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E peak [GB/s]
# Algorithmic bytes per particle of each reference stage (reference data flow, float4 = 16 B):
#   kick 48, drift 48, density 32, eos_micro 56, momentum 84, theta_update 36, kick 48,
#   sponge 8.8, rain 8, max_vel 16 (total 384.8); a re-sort adds 212 (keys 20, 3 radix passes
#   4 + 16*3, gather 108, ranges 32).
B_STAGE = dict(kick_drift=48 + 48, density_eos=32 + 56, momentum_kdk=84 + 36 + 48 + 8.8 + 16,
               rain=8.0)
B_STEP = 384.8
B_SORT = 212.0
# HBM bytes per particle per launch measured with FETCH_SIZE / WRITE_SIZE counters on this bench
# command (scaled by a 1 GiB device copy read back through the same counters), averaged over the
# 300 timed steps of the 256^3 workload; profiles/step_256cubed_hbm_traffic.md holds the table and
# the commands. Reported as roofline.traffic for the 256^3 workload only.
TRAFFIC_B_PER_PARTICLE = dict(momentum_kdk=198.2, density_eos=170.0, kick_drift=88.5, rain=24.6)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(cfg, st, dt, steps):
    """The FP32 oracle on the host cores: `steps` steps of the same workload."""
    from oracle.oracle import Oracle

    o = Oracle(cfg)
    o.set_state(st)
    o.sort()
    o.force(dt)
    t0 = time.perf_counter()
    for k in range(steps):
        o.step(dt, 0.2 * k, do_sort=(k == 0))
    wall = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=cfg.n * steps / wall, unit="particle-steps/s", cores=cores, kind="port",
                sample=f"{steps} steps (one with a re-sort) of the {cfg.nx}^3 workload, "
                       f"FP32 oracle, OpenMP, {wall:.1f} s")


def copy_probe(mib):
    """Three plain device-to-device copies of `mib` MiB through the HIP runtime's blit kernel: a
    kernel of known traffic (reads and writes mib MiB each) against which FETCH_SIZE and
    WRITE_SIZE counters are scaled."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = C.c_size_t(mib << 20)
    a, b = C.c_void_p(), C.c_void_p()
    ok = hip.hipMalloc(C.byref(a), n) == 0 and hip.hipMalloc(C.byref(b), n) == 0
    ok = ok and hip.hipMemset(a, 1, n) == 0 and hip.hipDeviceSynchronize() == 0
    for _ in range(3):
        ok = ok and hip.hipMemcpy(b, a, n, 3) == 0          # 3 = hipMemcpyDeviceToDevice
    ok = ok and hip.hipDeviceSynchronize() == 0
    for ptr in (a, b):
        if ptr.value:
            hip.hipFree(ptr)
    if not ok:
        raise SystemExit("copy probe: HIP runtime call failed")
    log(f"[bench] copy probe: 3 x {mib} MiB device-to-device")


def ensemble_bench(n_members, steps, seed=5, sequential=True):
    """The batched ensemble against sequential plain runs of the same members.

    Members from the library's ensemble preset (log-uniform 8,192-131,072 particles, 250 m
    lattice, moist, per-member noise seed), all re-sorted every step so both runs follow the
    same schedule. Batched: one handle, one launch per stage for all members, steps replayed
    from a captured hipGraph with no host synchronisation. Sequential: one handle per member,
    the reference loop (asph_run) with its per-step max|v| readback, default decomposition. After
    two untimed warm-up steps on both sides (the batched side autotunes its brick decomposition
    between them), `steps` steps are timed; the final states must agree bit for bit."""
    import numpy as np

    import sph_cuda_amd as s
    from sph_cuda_amd import _lib as L

    cfgs, ics, dt, _ = s.ensemble_preset(n_members, seed)
    for c in cfgs:
        c.flags |= L.F_SORT_EVERY_STEP
    states = [s.ic_generate(c, ic) for c, ic in zip(cfgs, ics)]
    ntot = sum(c.n for c in cfgs)
    log(f"[bench] ensemble: {n_members} members, {ntot} particles, dt {dt}")

    ens = s.Sim(cfgs)
    ens.set_state(s.concat_states(states))
    ens.run_graph(1, dt, sort_every=1)
    tune = ens.autotune(reps=2)
    log(f"[bench] ensemble autotune: {tune['best']} of {tune['trials']}")
    ens.run_graph(1, dt, sort_every=1)        # captures the step graph for the chosen shape
    ens.synchronize()
    t0 = time.perf_counter()
    ens.run_graph(steps, dt, sort_every=1)
    ens.synchronize()
    wall_b = time.perf_counter() - t0
    final = ens.get_state()
    log(f"[bench] batched: {steps} steps in {wall_b:.3f} s")

    wall_s, exact = 0.0, 0
    for m, (c, st) in enumerate(zip(cfgs, states) if sequential else []):
        with s.Sim(c) as one:
            one.set_state(st)
            rc, _ = one.run(max_steps=2, dt=dt, fixed_dt=True)
            one.synchronize()
            t0 = time.perf_counter()
            rc2, _ = one.run(max_steps=steps, dt=dt, fixed_dt=True)
            one.synchronize()
            wall_s += time.perf_counter() - t0
            if rc or rc2:
                raise SystemExit(f"member {m}: run aborted")
            o = one.get_state()
            a, n = ens.member_range(m)
            exact += int(all(np.array_equal(getattr(final, f)[a:a + n], getattr(o, f))
                             for f in ("pos", "vel", "thq", "rho")))
    log(f"[bench] sequential: {wall_s:.3f} s; members bit-identical: {exact}/{n_members}")
    return {
        "workload": f"{n_members}-member ensemble, {ntot} particles (8,192-131,072 per member, "
                    f"250 m lattice, nz 32), moist, dt {dt} s, re-sort every step",
        "members": n_members, "particles": ntot, "steps": steps,
        "value": ntot * steps / wall_b, "unit": "particle-steps/s",
        "ms_per_step": wall_b / steps * 1e3,
        "sequential_value": ntot * steps / wall_s if sequential else None,
        "sequential_ms_per_step": wall_s / steps * 1e3 if sequential else None,
        "speedup_vs_sequential": wall_s / wall_b if sequential else None,
        "members_bit_identical": exact if sequential else None,
        "tuning": tune["best"],
        "hbm_frac_step_alg": (B_STEP + B_SORT) * ntot * steps / wall_b / 1e9 / HBM_PEAK_GBS,
    }


def valu_probe():
    """The device's VALU issue roofline (asph_probe_valu): independent non-packed FP32 FMA chains.
    The neighbour kernels' own instruction counts come from SQ counters (profiles/)."""
    import ctypes as C

    from sph_cuda_amd import _lib as L

    w = C.c_double()
    L.check(L.load().asph_probe_valu(8192, 5, C.byref(w)), "asph_probe_valu")
    return {"wave_fma_per_s": w.value, "fp32_nonpacked_tflops": w.value * 128 / 1e12,
            "wave_fma_per_simd_cycle_at_2p4GHz": w.value / (2.4e9 * 1024)}


def device_loop_bench(steps=120, poll=32, reps=3):
    """The reference loop at the 32^3 dry workload (preset 1, CFL-controlled dt): the host loop
    (asph_run, one host synchronisation per step for max|v|) against the same loop with its
    controller on the device (asph_run_device, a captured graph and a poll every `poll` steps),
    each after the autotuner. Best of `reps`; the final states must agree bit for bit."""
    import numpy as np

    import sph_cuda_amd as s

    cfg, ic, dt, _ = s.preset(1)
    st = s.ic_generate(cfg, ic)
    walls = {"host": None, "device": None}
    finals = {}
    for name in ("host", "device"):
        for _ in range(reps):
            with s.Sim(cfg) as g:
                g.set_state(st)
                g.sort()
                g.force(dt)
                g.autotune(reps=2)
                g.synchronize()
                t0 = time.perf_counter()
                if name == "host":
                    rc, stats = g.run(max_steps=steps)
                else:
                    rc, stats = g.run_device(max_steps=steps, poll=poll)
                g.synchronize()
                w = time.perf_counter() - t0
                if rc != 0:
                    raise SystemExit(f"device-loop bench: {name} run aborted")
                walls[name] = w if walls[name] is None else min(walls[name], w)
                finals[name] = g.get_state()
    same = all(np.array_equal(getattr(finals["host"], f), getattr(finals["device"], f))
               for f in ("pos", "vel", "thq", "rho"))
    n = cfg.n
    log(f"[bench] device loop: host {walls['host'] * 1e3:.1f} ms, device {walls['device'] * 1e3:.1f} ms, same {same}")
    return {
        "workload": f"dry bubble {cfg.nx}x{cfg.ny}x{cfg.nz} = {n} particles, {steps} CFL-controlled steps",
        "host_loop_value": n * steps / walls["host"], "device_loop_value": n * steps / walls["device"],
        "unit": "particle-steps/s", "host_loop_ms_per_step": walls["host"] / steps * 1e3,
        "device_loop_ms_per_step": walls["device"] / steps * 1e3, "speedup": walls["host"] / walls["device"],
        "poll_steps": poll, "bit_identical": bool(same),
    }


def adjoint_bench(reps=3, nsteps=60, checkpoint=10):
    """The adjoint of column-integrated rain at the 64^3 adjoint workload (preset 4): one
    asph_adjoint_rain call (forward window with checkpoints, recomputation and reverse sweep)
    against a plain forward of the same window and sort schedule. Best of `reps` of each."""
    import sph_cuda_amd as s

    cfg, ic, dt, _ = s.preset(4)
    st = s.ic_generate(cfg, ic)
    sim = s.Sim(cfg)
    sim.set_state(st)
    sim.adjoint_rain(2, dt, checkpoint)   # warm-up (allocations, code paths)
    best = None
    for _ in range(reps):
        r = sim.adjoint_rain(nsteps, dt, checkpoint)
        if best is None or r["forward_s"] + r["adjoint_s"] < best["forward_s"] + best["adjoint_s"]:
            best = r
    plain = None
    for _ in range(reps):
        f = s.Sim(cfg)
        f.set_state(st)
        f.synchronize()
        t0 = time.perf_counter()
        f.sort()
        f.force(dt)
        for k in range(nsteps):
            f.step(dt, k * dt, do_sort=(k > 0 and k % checkpoint == 0))
        f.synchronize()
        w = time.perf_counter() - t0
        plain = w if plain is None else min(plain, w)
        del f
    total = best["forward_s"] + best["adjoint_s"]
    log(f"[bench] adjoint: value+gradient {total * 1e3:.1f} ms, plain forward {plain * 1e3:.1f} ms")
    return {
        "workload": f"rain adjoint, {cfg.nx}x{cfg.ny}x{cfg.nz} = {cfg.n} particles, {nsteps} steps of "
                    f"dt={dt} s, checkpoint every {checkpoint}",
        "plain_forward_ms": plain * 1e3,
        "adjoint_call_ms": total * 1e3,
        "adjoint_forward_pass_ms": best["forward_s"] * 1e3,
        "adjoint_reverse_ms": best["adjoint_s"] * 1e3,
        "ratio": (total - plain) / plain,
        "ratio_definition": "(value+gradient call - plain forward) / plain forward: the reverse "
                            "sweep, the checkpoint recomputation and the taping, over the forward",
        "segments": int(best["segments"]),
        "J_kg": best["J"],
        "dJ_dqc0": best["dJ_dqc0"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--preset", type=int, default=3)
    ap.add_argument("--n", type=int, default=0, help="override lattice edge (cubic)")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=240)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the batched leg of the ensemble (profiling)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-device-loop", action="store_true", help="skip the device-controller leg")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint leg (profiling)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the run, copy a buffer of this size 3 times (device to device) as "
                         "the byte reference for hardware traffic counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the step is a single-GPU kernel set "
              "(multi-GPU slab decomposition is a design, not an implementation)", file=sys.stderr)
        sys.exit(2)

    import sph_cuda_amd as s

    if args.ensemble_only:
        print(json.dumps(ensemble_bench(args.ensemble_members, args.ensemble_steps, sequential=False)),
              flush=True)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_bench()), flush=True)
        return
    cfg, ic, dt, _ = s.preset(args.preset)
    if args.n:
        cfg.nx = cfg.ny = cfg.nz = args.n
    log(f"[bench] preset {args.preset}: {cfg.nx}x{cfg.ny}x{cfg.nz} = {cfg.n} particles, dt {dt}")
    t0 = time.perf_counter()
    st = s.ic_generate(cfg, ic)
    log(f"[bench] initial state generated in {time.perf_counter() - t0:.1f} s")

    sim = s.Sim(cfg)
    sim.set_state(st)
    rc, ws = sim.run(max_steps=args.warmup, dt=dt, fixed_dt=True)
    if rc != 0:
        raise SystemExit(f"warmup aborted rc={rc}")
    sim.synchronize()
    log(f"[bench] warmup {args.warmup} steps done ({ws.sorts} sorts)")
    tune = sim.autotune(reps=2)
    log(f"[bench] autotune: {tune['best']} of {tune['trials']}")

    sim.set_profiling(True)
    sim.kernel_times()   # reset
    sim.synchronize()
    t0 = time.perf_counter()
    rc, ts = sim.run(max_steps=args.steps, dt=dt, fixed_dt=True)
    sim.synchronize()
    wall = time.perf_counter() - t0
    if rc != 0:
        raise SystemExit(f"timed run aborted rc={rc}")
    kt = sim.kernel_times()
    log(f"[bench] timed {args.steps} steps in {wall:.3f} s, {ts.sorts} sorts; kernels: {kt}")

    n = cfg.n
    value = n * args.steps / wall
    # dominant kernel by total time
    dom = max(("kick_drift", "density_eos", "momentum_kdk", "rain"), key=lambda k: kt[k][0])
    ms_total, cnt = kt[dom]
    avg_s = ms_total / 1e3 / max(cnt, 1)
    bytes_launch = B_STAGE[dom] * n
    achieved = bytes_launch / avg_s / 1e9
    traffic = None
    if args.preset == 3 and not args.n and args.steps == 300 and args.warmup == 10:
        traffic = TRAFFIC_B_PER_PARTICLE[dom] * n
    step_alg_gbs = (B_STEP * n * args.steps + B_SORT * n * ts.sorts) / wall / 1e9

    out = {
        "metric": "particle-steps/s & % HBM peak (plain vs batched); adjoint/forward time ratio",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded initial state, no dataset)",
        "config": {
            "workload": f"plain moist anisotropic step, {cfg.nx}x{cfg.ny}x{cfg.nz} = {n} particles, "
                        f"dx=dy={cfg.dx} m dz={cfg.dz} m, h=1.3 cells, dt={dt} s fixed",
            "particles": n, "sorts_in_timed_steps": int(ts.sorts),
            "hbm_frac_step_alg": step_alg_gbs / HBM_PEAK_GBS,
            "step_alg_bytes_per_particle": B_STEP, "sort_alg_bytes_per_particle": B_SORT,
            "kernel_ms_per_launch": {k: (v[0] / max(v[1], 1)) for k, v in kt.items()},
            "kernel_launches": {k: v[1] for k, v in kt.items()},
            "parallelism": "single GPU",
            "tuning": tune["best"],
        },
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "alg_bytes_per_particle": B_STAGE[dom], "traffic": traffic,
            "traffic_GBps_at_measured_time": (traffic / avg_s / 1e9) if traffic else None,
        },
        "cpu_baseline": None,
        "batched": None,
        "adjoint": None,
        "device_loop": None,
    }
    del sim
    if not args.no_ensemble:
        out["batched"] = ensemble_bench(args.ensemble_members, args.ensemble_steps)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
    if not args.no_device_loop:
        out["device_loop"] = device_loop_bench()
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
    out["valu_probe"] = valu_probe()
    if not args.no_cpu_baseline:
        log("[bench] CPU oracle sample ...")
        out["cpu_baseline"] = cpu_baseline(cfg, st, dt, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
