"""GPU parity at the workloads' own sizes, against the CPU oracle (OpenMP) on the same inputs:
- the reference's 32^3 dry warm bubble run in full (120 steps under the CFL controller, with its
  displacement-gated sorts);
- the 128^3 moist workload and the 256^3 headline workload: the initial sort, the force
  evaluation and one re-sorted KDK step;
- the 256^3 sort also against the closed form (stable argsort of the Morton keys, cell ranges
  from the sorted keys).

Bars as in the small-size parity tests: sort arrays and the rain grid bit-exact, the state within
1e-5 normwise per field (the kernels evaluate the oracle's canonical arithmetic; the printed
maximum differences show whether they agree to the bit).
"""
import numpy as np
import pytest

import sph_cuda_amd as s
from helpers import state_close
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5
FIELDS = ("pos", "vel", "thq", "rho")


def _max_abs(a, b):
    return {f: float(np.max(np.abs(getattr(a, f).astype(np.float64) - getattr(b, f)))) for f in FIELDS}


def _pair(which):
    cfg, ic, dt, steps = s.preset(which)
    st = s.ic_generate(cfg, ic)
    g, o = s.Sim(cfg), Oracle(cfg)
    g.set_state(st)
    o.set_state(st)
    print(f"[case] preset {which}: {cfg.nx}x{cfg.ny}x{cfg.nz} = {cfg.n} particles, dt {dt}", flush=True)
    return cfg, dt, steps, g, o


def test_reference_case_full_run():
    cfg, dt, steps, g, o = _pair(1)
    rc_g, sg = g.run(max_steps=steps, dt=dt)
    rc_o, so = o.run(max_steps=steps, dt=dt)
    print(f"[case] device steps {sg.steps} sorts {sg.sorts} trips {sg.trips}; oracle {so}", flush=True)
    assert rc_g == rc_o == 0
    assert sg.steps == so["steps"] == steps and sg.sorts == so["sorts"] and sg.trips == so["trips"]
    a, b = g.get_state(), o.get_state()
    bad, errs = state_close(a, b, TOL)
    print(f"[errs] {errs}; max |diff| {_max_abs(a, b)}", flush=True)
    assert not bad, errs
    assert np.array_equal(g.get_rain_fx()[0], o.get_rain_fx()[0])


@pytest.mark.parametrize("which", [2, 3])
def test_one_step_at_full_size(which):
    cfg, dt, _, g, o = _pair(which)
    for x in (g, o):
        x.sort()
    a, b = g.get_sort(), o.get_sort()
    for k in ("keys", "keys_sorted", "perm", "cell_range"):
        assert np.array_equal(a[k], b[k]), k
    if which == 3:
        # the closed form of the sort: stable argsort of the keys, ranges from the sorted keys
        perm = np.argsort(a["keys"], kind="stable").astype(np.uint32)
        assert np.array_equal(a["perm"], perm)
        ks = a["keys"][perm]
        assert np.array_equal(a["keys_sorted"], ks)
        cr = a["cell_range"]
        nonempty = cr[:, 1] > cr[:, 0]
        assert int((cr[:, 1] - cr[:, 0]).sum()) == cfg.n
        assert np.array_equal(np.flatnonzero(np.r_[True, ks[1:] != ks[:-1]]), np.sort(cr[nonempty, 0]))
    for x in (g, o):
        x.force(dt)
    ga, gp, gd = g.get_aux()
    oa, op, od = o.get_aux()
    assert np.array_equal(g.get_state().rho, o.get_state().rho)
    print(f"[case] force: max |acc diff| {float(np.max(np.abs(ga - oa))):.3e}", flush=True)
    for x in (g, o):
        x.step(dt, 0.0, do_sort=True)
    sa, sb = g.get_state(), o.get_state()
    bad, errs = state_close(sa, sb, TOL)
    print(f"[errs] {errs}; max |diff| {_max_abs(sa, sb)}", flush=True)
    assert not bad, errs
    assert np.array_equal(g.get_rain_fx()[0], o.get_rain_fx()[0])
    assert np.array_equal(g.max_vel(), o.max_vel())


@pytest.mark.parametrize("which,steps", [(2, 24), (3, 4)])
def test_controlled_run_at_full_size(which, steps):
    """The reference loop under the CFL controller (displacement-gated sorts, dt capped at the
    workload's) from t = 0 at the workload's own size: the 128^3 workload over 24 steps and the
    256^3 workload over 4, state, rain grid and loop statistics bit for bit."""
    cfg, dt, _, g, o = _pair(which)
    rc_g, sg = g.run(max_steps=steps)
    rc_o, so = o.run(max_steps=steps)
    assert rc_g == rc_o == 0
    assert (sg.steps, sg.sorts, sg.trips) == (so["steps"], so["sorts"], so["trips"])
    assert sg.simt == so["simt"] and np.float32(sg.dt) == np.float32(so["dt"])
    a, b = g.get_state(), o.get_state()
    print(f"[case] preset {which}: {steps} steps, {sg.sorts} sorts, t = {sg.simt:.2f} s; max |diff| "
          f"{_max_abs(a, b)}", flush=True)
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(g.get_rain_fx()[0], o.get_rain_fx()[0])
