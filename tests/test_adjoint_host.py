"""The FP64 host adjoint (oracle/asph_adjoint_ref.cpp) against central differences of the FP64
oracle: pins the derivative formulas that the device adjoint shares (pointwise stages, pair
terms, sedimentation transpose, kick/drift with wall reflection, segment sorts)."""
import numpy as np
import pytest

import sph_cuda_amd as s
from helpers import small_config
from oracle.oracle import adjoint_rain_f64, rain_objective_f64

NSTEPS, CK = 6, 3


@pytest.fixture(scope="module")
def case():
    cfg, ic, dt = small_config("adj", (12, 12, 10))
    st = s.ic_generate(cfg, ic)
    J, g, gq = adjoint_rain_f64(cfg, st, NSTEPS, dt, CK)
    return cfg, dt, st, J, g, gq


def _J(cfg, dt, st, thq):
    return rain_objective_f64(cfg, st.pos, st.vel, thq, st.rho, NSTEPS, dt, CK)


def test_objective_matches_forward(case):
    cfg, dt, st, J, g, gq = case
    assert J > 0 and abs(J - _J(cfg, dt, st, st.thq.astype(np.float64))) <= 1e-12 * J


@pytest.mark.parametrize("comp,eps", [(0, 1e-4), (1, 1e-8), (2, 1e-8)])
def test_state_gradient_matches_central_differences(case, comp, eps):
    cfg, dt, st, J, g, gq = case
    rng = np.random.default_rng(4000 + comp)
    base = st.thq.astype(np.float64)
    for k in range(3):
        d = rng.standard_normal(cfg.n)
        tp, tm = base.copy(), base.copy()
        tp[:, comp] += eps * d
        tm[:, comp] -= eps * d
        fd = (_J(cfg, dt, st, tp) - _J(cfg, dt, st, tm)) / (2 * eps)
        ad = float(np.dot(g[:, comp], d))
        scale = np.linalg.norm(g[:, comp]) * np.linalg.norm(d)
        print(f"[case] comp {comp} dir {k}: FD {fd:.9e} adjoint {ad:.9e}", flush=True)
        # relative to the derivative, or to 1% of |g||d| for a direction nearly orthogonal to g
        assert abs(ad - fd) <= 2e-5 * max(abs(fd), 1e-2 * scale)


def test_threshold_gradient_matches_central_differences(case):
    cfg, dt, st, J, g, gq = case
    # a small step: a larger one moves the autoconversion switch qc > qc0 across particles whose
    # cloud water lies within it (a kink of the model; 1e-5 and 1e-6 are off by ~2%)
    e = 1e-8
    cp, cm = cfg.copy(), cfg.copy()
    cp.qc_auto, cm.qc_auto = cfg.qc_auto + e, cfg.qc_auto - e
    de = (float(np.float32(cp.qc_auto)) - float(np.float32(cm.qc_auto))) / 2
    th = st.thq.astype(np.float64)
    fd = (rain_objective_f64(cp, st.pos, st.vel, th, st.rho, NSTEPS, dt, CK)
          - rain_objective_f64(cm, st.pos, st.vel, th, st.rho, NSTEPS, dt, CK)) / (2 * de)
    print(f"[case] qc0: FD {fd:.9e} adjoint {gq:.9e}", flush=True)
    assert gq < 0 and abs(gq - fd) <= 2e-6 * abs(fd)


def _windy(cfg, st, cells, nsteps, dt):
    """The state with a uniform horizontal wind that carries every particle `cells` cells along x
    (and half of that along y) over the window: listed and current cells drift apart by more than
    one cell between sorts, the regime of the forward's global-table walk."""
    vel = st.vel.copy()
    vel[:, 0] += np.float32(cells * cfg.dx / (nsteps * dt))
    vel[:, 1] += np.float32(0.5 * cells * cfg.dy / (nsteps * dt))
    return s.State(st.pos, vel, st.thq, st.rho)


def test_pair_enumeration_is_complete_with_stale_lists():
    """With drift beyond a cell between sorts, widening the pair walk by 3 cells per axis (the whole
    x-y lattice here) finds no further pair: the gradient changes only by summation order."""
    cfg, ic, dt = small_config("adj", (12, 12, 10))
    n, ck = 6, 6
    st = _windy(cfg, s.ic_generate(cfg, ic), 2.5, n, dt)
    J, g, gq = adjoint_rain_f64(cfg, st, n, dt, ck)
    J2, g2, gq2 = adjoint_rain_f64(cfg, st, n, dt, ck, extra_margin=3)
    assert J == J2
    assert np.max(np.abs(g - g2)) <= 1e-12 * np.max(np.abs(g))
    assert abs(gq - gq2) <= 1e-12 * abs(gq)


def test_gradient_with_stale_lists_beyond_one_cell():
    """Finite differences in that regime. A pair enters or leaves the stencil when a particle crosses
    a cell face, so J has small jumps there and a central difference carries O(jump / eps) noise
    (3e-5 relative at eps 1e-5 to 1e-4 on this case, against 2e-7 without wind): the bound is wider."""
    cfg, ic, dt = small_config("adj", (12, 12, 10))
    n, ck = 6, 6   # one segment: no re-sort inside the window
    st = _windy(cfg, s.ic_generate(cfg, ic), 2.5, n, dt)
    J, g, _ = adjoint_rain_f64(cfg, st, n, dt, ck)
    rng = np.random.default_rng(4100)
    base = st.thq.astype(np.float64)
    eps = 1e-4
    for k in range(2):
        d = rng.standard_normal(cfg.n)
        tp, tm = base.copy(), base.copy()
        tp[:, 0] += eps * d
        tm[:, 0] -= eps * d
        fd = (rain_objective_f64(cfg, st.pos, st.vel, tp, st.rho, n, dt, ck)
              - rain_objective_f64(cfg, st.pos, st.vel, tm, st.rho, n, dt, ck)) / (2 * eps)
        ad = float(np.dot(g[:, 0], d))
        scale = np.linalg.norm(g[:, 0]) * np.linalg.norm(d)
        print(f"[case] windy dir {k}: FD {fd:.9e} adjoint {ad:.9e}", flush=True)
        assert abs(ad - fd) <= 2e-4 * max(abs(fd), 1e-2 * scale)


def test_fused_momentum_pair_matches_two_orientations():
    """The device's fused evaluation of both orientations of a momentum pair (mom_pair_vjp2)
    against two evaluations of the single-orientation terms, on random pairs in the support (a
    quarter with only one orientation seeded): equal up to rounding in double; in float, within
    the cancellation of the summed terms."""
    import ctypes as C
    from oracle.oracle import load
    cfg, ic, dt = small_config("adj", (12, 12, 10))
    lib = load()
    assert lib.oracle_check_mom_pair_fused(C.byref(cfg), 20000, 7, 0) <= 1e-12
    assert lib.oracle_check_mom_pair_fused(C.byref(cfg), 20000, 7, 1) <= 1e-4


@pytest.mark.parametrize("kind", ["adj", "moist"])
def test_reverse_mode_momentum_epilogue_matches_forward_mode(kind):
    """The device's reverse-mode product of the momentum epilogue (mom_vjp_rev: analytic terms and
    one dual pass in z) against the forward-mode product (one dual pass per input), on random
    inputs that include the clamps of dtheta/dt, of theta and of the pressure and the sponge layer:
    equal up to rounding in double, within float rounding in float."""
    import ctypes as C
    from oracle.oracle import load
    cfg, ic, dt = small_config(kind, (12, 12, 10))
    lib = load()
    for mode in (0, 1):
        assert lib.oracle_check_mom_vjp_rev(C.byref(cfg), 20000, 11, 0, mode) <= 1e-13
        assert lib.oracle_check_mom_vjp_rev(C.byref(cfg), 20000, 11, 1, mode) <= 1e-5
