"""GPU against the CPU oracle on the workload paths that used to be checked only against other GPU
paths:
- the clumped state the 256^3 physics reaches by itself (a 16 x 16 x 256 column of that workload
  under the reference's CFL controller, asph_h100.cu:1040-1078, for 450 steps: the rigid lid
  pushes its top layers down and the particles clump, so the neighbour kernels take their slabbed
  paths with several times the initial pair count);
- the device-resident loop controller (asph_run_device) against the oracle's own run loop;
- ensemble members and the 256-member ensemble preset under graph replay, each member against its
  own oracle.

Bars as elsewhere: the state bit for bit (the kernels evaluate the oracle's canonical arithmetic),
and the rain grid exactly.
"""
import time

import numpy as np
import pytest

import sph_cuda_amd as s
from helpers import small_config
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "thq", "rho")


def _same(a, b, sl=slice(None)):
    return {f: bool(np.array_equal(getattr(a, f)[sl], getattr(b, f))) for f in FIELDS}


def test_clumped_column_matches_oracle():
    cfg, ic, dt, _ = s.preset(3)
    cfg.nx = cfg.ny = 16
    cfg.dt_cap = 0.2
    st = s.ic_generate(cfg, ic)
    print(f"[case] column {cfg.nx}x{cfg.ny}x{cfg.nz}, 256^3 physics, 450 controlled steps", flush=True)
    g = s.Sim(cfg)
    g.set_state(st)
    g.set_counters(True)
    rc1, s1 = g.run(max_steps=50, dt=0.2)
    c0 = g.counters()
    rc2, s2 = g.run(max_steps=350)
    g.counters()
    rc3, s3 = g.run(max_steps=50)
    c1 = g.counters()
    assert rc1 == rc2 == rc3 == 0
    early, late = c0["mom_pairs"] / (50 * cfg.n), c1["mom_pairs"] / (50 * cfg.n)
    print(f"[case] pairs per particle {early:.1f} -> {late:.1f}; late slabbed bricks "
          f"{c1['mom_slabbed']}, one-slab {c1['mom_one_slab']}, global {c1['mom_global']}, "
          f"mode B {c1['mom_modeb']}; |w| max {s3.umax[2]:.1f} m/s, dt {s3.dt:.4f}", flush=True)
    assert late > 2.0 * early and c1["mom_slabbed"] > 0
    t0 = time.perf_counter()
    o = Oracle(cfg)
    o.set_state(st)
    ro = [o.run(max_steps=50, dt=0.2), o.run(max_steps=350), o.run(max_steps=50)]
    print(f"[case] oracle 450 steps in {time.perf_counter() - t0:.1f} s", flush=True)
    for (rc, so), sg in zip(ro, (s1, s2, s3)):
        assert rc == 0
        assert (so["steps"], so["sorts"], so["trips"]) == (sg.steps, sg.sorts, sg.trips)
        assert so["simt"] == sg.simt and np.float32(so["dt"]) == np.float32(sg.dt)
    same = _same(o.get_state(), g.get_state())
    assert all(same.values()), same
    rg, pg, cg = g.get_rain_fx()
    r_o, p_o, c_o = o.get_rain_fx()
    assert np.array_equal(rg, r_o) and np.array_equal(pg, p_o) and cg == c_o


@pytest.mark.parametrize("kind,steps,poll", [("moist", 13, 4), ("ref", 30, 8), ("dry", 7, 16)])
def test_device_loop_matches_oracle_run(kind, steps, poll):
    cfg, ic, dt = small_config(kind, (16, 16, 12))
    st0 = s.ic_generate(cfg, ic)
    print(f"[case] {kind} steps {steps} poll {poll}", flush=True)
    with s.Sim(cfg) as d:
        d.set_state(st0)
        o = Oracle(cfg)
        o.set_state(st0)
        for n in (steps, 5):
            rc_d, sd = d.run_device(max_steps=n, poll=poll)
            rc_o, so = o.run(max_steps=n)
            assert rc_d == rc_o == 0
            assert (sd.steps, sd.sorts, sd.trips) == (so["steps"], so["sorts"], so["trips"])
            assert sd.simt == so["simt"] and np.float32(sd.dt) == np.float32(so["dt"])
            same = _same(o.get_state(), d.get_state())
            assert all(same.values()), same


@pytest.mark.parametrize("strict", [1, 0])
def test_device_loop_cfl_abort_and_clamp_match_oracle(strict):
    cfg, ic, dt = small_config("ref")
    cfg.strict_cfl = strict
    st0 = s.ic_generate(cfg, ic)
    st0.vel[:, 0] = 20000.0
    with s.Sim(cfg) as d:
        d.set_state(st0)
        o = Oracle(cfg)
        o.set_state(st0)
        rc_d, sd = d.run_device(max_steps=6, poll=4)
        rc_o, so = o.run(max_steps=6)
        print(f"[case] strict {strict}: rc {rc_d} {rc_o} steps {sd.steps} {so['steps']} trips {sd.trips} "
              f"{so['trips']}", flush=True)
        assert rc_d == rc_o and (sd.steps, sd.trips) == (so["steps"], so["trips"]) and sd.trips >= 1
        same = _same(o.get_state(), d.get_state())
        assert all(same.values()), same


# members of different shapes; the 12 x 12 x 10 one is narrower than a brick halo (thread-per-
# particle path), the others take the LDS brick path
SHAPES = [(16, 16, 12), (20, 18, 12), (12, 12, 10), (24, 24, 14)]


def test_ensemble_members_match_oracle():
    mem = []
    for k, sh in enumerate(SHAPES):
        cfg, ic, dt = small_config("moist", sh)
        ic.seed = 100 + k
        ic.white_noise = 0.1
        mem.append((cfg, ic))
    states = [s.ic_generate(c, ic) for c, ic in mem]
    ens = s.Sim([c for c, _ in mem])
    ens.set_state(s.concat_states(states))
    ens.sort()
    ens.force(dt)
    for k in range(6):
        ens.step(dt, k * dt, do_sort=(k % 2 == 0))
    e_state = ens.get_state()
    e_rain, e_prec, _ = ens.get_rain_fx()
    col = 0
    for m, (c, _) in enumerate(mem):
        o = Oracle(c)
        o.set_state(states[m])
        o.sort()
        o.force(dt)
        for k in range(6):
            o.step(dt, k * dt, do_sort=(k % 2 == 0))
        a, n = ens.member_range(m)
        same = _same(e_state, o.get_state(), slice(a, a + n))
        print(f"[case] member {m} shape {SHAPES[m]}: {same}", flush=True)
        assert all(same.values()), (m, same)
        r, p, _ = o.get_rain_fx()
        assert np.array_equal(e_rain[a:a + n], r)
        assert np.array_equal(e_prec[col:col + c.nx * c.ny], p)
        col += c.nx * c.ny
        assert np.array_equal(ens.member_max_vel(m), o.max_vel())


def test_ensemble_preset_256_members_graph_vs_oracle():
    """The batched workload itself: 256 members (log-uniform 8,192-131,072 particles), two steps
    replayed from the captured graph with a sort every step, every member against its own FP32
    oracle."""
    cfgs, ics, dt, _ = s.ensemble_preset(256, 5)
    for c in cfgs:
        c.flags |= s._lib.F_SORT_EVERY_STEP
    states = [s.ic_generate(c, ic) for c, ic in zip(cfgs, ics)]
    ntot = sum(c.n for c in cfgs)
    print(f"[case] ensemble preset: {len(cfgs)} members, {ntot} particles", flush=True)
    ens = s.Sim(cfgs)
    ens.set_state(s.concat_states(states))
    st = ens.run_graph(2, dt, sort_every=1)
    assert st.steps == 2 and st.sorts == 2
    final = ens.get_state()
    rain, prec, _ = ens.get_rain_fx()
    t0 = time.perf_counter()
    col, bad = 0, []
    for m, (c, x) in enumerate(zip(cfgs, states)):
        o = Oracle(c)
        o.set_state(x)
        o.sort()
        o.force(dt)
        for k in range(2):
            o.step(dt, k * dt, do_sort=True)
        a, n = ens.member_range(m)
        same = _same(final, o.get_state(), slice(a, a + n))
        r, p, _ = o.get_rain_fx()
        ok = all(same.values()) and np.array_equal(rain[a:a + n], r) and \
            np.array_equal(prec[col:col + c.nx * c.ny], p)
        col += c.nx * c.ny
        if not ok:
            bad.append((m, same))
    print(f"[case] oracle of 256 members in {time.perf_counter() - t0:.1f} s; mismatched {bad[:4]}",
          flush=True)
    assert not bad
