"""The reference loop with its controller on the device (asph_run_device) against the host loop
(asph_run): the same steps, sorts, CFL decisions, dumps and final state, bit for bit."""
import os

import numpy as np
import pytest

import sph_cuda_amd as s
from helpers import small_config

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "thq", "rho")


def _same_state(a, b):
    return all(np.array_equal(getattr(a, f), getattr(b, f)) for f in FIELDS)


def _stats(st):
    return (st.steps, st.sorts, st.trips, st.dumps, st.simt, st.dt, st.dt_required, tuple(st.umax))


@pytest.mark.parametrize("kind,steps,poll", [("moist", 13, 4), ("ref", 30, 8), ("dry", 7, 16)])
def test_adaptive_dt_matches_host_loop(kind, steps, poll):
    cfg, ic, dt = small_config(kind, (16, 16, 12))
    st0 = s.ic_generate(cfg, ic)
    print(f"[case] {kind} steps {steps} poll {poll}", flush=True)
    with s.Sim(cfg) as g, s.Sim(cfg) as d:
        g.set_state(st0)
        d.set_state(st0)
        rc1, s1 = g.run(max_steps=steps)
        rc2, s2 = d.run_device(max_steps=steps, poll=poll)
        assert rc1 == rc2 == 0
        assert _stats(s1)[:3] == _stats(s2)[:3] and _stats(s1)[4:] == _stats(s2)[4:], (_stats(s1), _stats(s2))
        assert _same_state(g.get_state(), d.get_state())
        # and a second run continues from the same loop state (sort interval, time since the sort)
        rc1, s1 = g.run(max_steps=5)
        rc2, s2 = d.run_device(max_steps=5, poll=poll)
        assert _stats(s1)[:3] == _stats(s2)[:3] and _same_state(g.get_state(), d.get_state())


def test_sim_end_and_dumps_match_host_loop(tmp_path):
    cfg, ic, dt = small_config("moist", (16, 16, 12))
    st0 = s.ic_generate(cfg, ic)
    o1, o2 = str(tmp_path / "host"), str(tmp_path / "device")
    os.makedirs(o1)
    os.makedirs(o2)
    with s.Sim(cfg) as g, s.Sim(cfg) as d:
        g.set_state(st0)
        d.set_state(st0)
        rc1, s1 = g.run(sim_end=12.0, dump_every=2.5, dump_stride=(4, 4, 3), outdir=o1, validate=True)
        rc2, s2 = d.run_device(sim_end=12.0, dump_every=2.5, dump_stride=(4, 4, 3), outdir=o2, validate=True,
                               poll=6)
        print(f"[case] steps {s1.steps} {s2.steps} dumps {s1.dumps} {s2.dumps}", flush=True)
        assert rc1 == rc2 == 0 and s1.steps == s2.steps and s1.dumps == s2.dumps >= 3
        assert s1.simt == s2.simt and _same_state(g.get_state(), d.get_state())
    f1, f2 = sorted(os.listdir(o1)), sorted(os.listdir(o2))
    assert f1 == f2
    for f in f1:
        with open(os.path.join(o1, f), "rb") as a, open(os.path.join(o2, f), "rb") as b:
            assert a.read() == b.read(), f


@pytest.mark.parametrize("strict", [1, 0])
def test_cfl_trip_clamp_and_abort_match_host_loop(strict):
    """|u| = 20 km/s: the required dt falls below the floor (asph_h100.cu:1054-1071). Strict mode
    aborts after the first step; otherwise the velocities are clamped and the run goes on."""
    cfg, ic, dt = small_config("ref")
    cfg.strict_cfl = strict
    st0 = s.ic_generate(cfg, ic)
    st0.vel[:, 0] = 20000.0
    with s.Sim(cfg) as g, s.Sim(cfg) as d:
        g.set_state(st0)
        d.set_state(st0)
        rc1, s1 = g.run(max_steps=6)
        rc2, s2 = d.run_device(max_steps=6, poll=4)
        print(f"[case] strict {strict}: rc {rc1} {rc2} steps {s1.steps} {s2.steps} trips {s1.trips} {s2.trips}",
              flush=True)
        assert rc1 == rc2 and s1.steps == s2.steps and s1.trips == s2.trips and s1.trips >= 1
        assert rc1 == (s._lib.CFL_ABORT if strict else 0)
        assert _same_state(g.get_state(), d.get_state())


def test_run_without_a_stop_condition_is_rejected():
    """Neither sim_end nor max_steps: both loops return an error instead of running forever."""
    cfg, ic, dt = small_config("ref")
    with s.Sim(cfg) as g:
        g.set_state(s.ic_generate(cfg, ic))
        with pytest.raises(s.AsphError, match="sim_end or max_steps"):
            g.run()
        with pytest.raises(s.AsphError, match="sim_end or max_steps"):
            g.run_device()
