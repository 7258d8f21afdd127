"""Pins of the CPU oracle against closed forms, invariants and known answers (no GPU).

Each test names the reference code whose behaviour it pins (asph_h100.cu line numbers)."""
import math

import numpy as np
import pytest

import sph_cuda_amd as s
from helpers import small_config
from oracle.oracle import Oracle, load, scalar

# the reference's float constants (asph_h100.cu:66-70) as exact doubles
GRAV, R_D, CP, P0, L_V = (float(np.float32(v)) for v in (9.80665, 287.05, 1004.0, 1.0e5, 2.501e6))


def test_morton_known_answers():
    """part1by2_10 / morton3, asph_h100.cu:176-186."""
    lib = load()
    assert lib.oracle_morton3(0, 0, 0) == 0
    assert lib.oracle_morton3(1, 0, 0) == 1
    assert lib.oracle_morton3(0, 1, 0) == 2
    assert lib.oracle_morton3(0, 0, 1) == 4
    assert lib.oracle_morton3(1023, 1023, 1023) == 0x3FFFFFFF
    assert lib.oracle_morton3(2, 0, 0) == 8
    assert lib.oracle_morton3(3, 5, 6) == 0b110101011  # z2y2x2? bits: x=011,y=101,z=110
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(0, 1024, 3))
        k = 0
        for b in range(10):
            k |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        assert lib.oracle_morton3(x, y, z) == k


@pytest.mark.parametrize("n,bits", [((32, 32, 32), 15), ((128, 128, 128), 21),
                                    ((256, 256, 256), 24), ((64, 64, 64), 18),
                                    ((1000, 1000, 150), 30), ((128, 128, 64), 21)])
def test_sort_end_bit(n, bits):
    """endbit = 3 * max ceil_log2 (asph_h100.cu:866-868)."""
    cfg = s.reference_config(*n)
    assert Oracle(cfg).end_bit() == bits


def test_spline_closed_forms():
    """M4 spline (asph_h100.cu:125-134): normalised to 1 in 3-D with sigma = 1/pi, continuous
    at q = 1 and 2, and dW/dq / q = kgrad (h = 1)."""
    cfg = s.reference_config(8, 8, 8)
    sigma = float(np.float32(0.31830988618))   # SIGMA as the float constant of :56
    f = lambda q: scalar(cfg, "fker", q, fp64=True)
    k = lambda q: scalar(cfg, "kgrad", q, fp64=True)
    q = np.linspace(0, 2, 20001)
    w = np.array([f(x) for x in q])
    integral = np.trapezoid(4 * np.pi * q**2 * sigma * w, q)
    assert abs(integral - 1.0) < 1e-6
    assert abs(f(1 - 1e-12) - f(1 + 1e-12)) < 1e-10 and abs(f(1.0) - 0.25) < 1e-12
    assert f(2.0) == 0.0 and f(2.5) == 0.0 and f(0.0) == 1.0
    for x in (0.1, 0.5, 0.9, 1.1, 1.5, 1.9):
        e = 1e-6
        dwdq = sigma * (f(x + e) - f(x - e)) / (2 * e)
        assert abs(k(x) * x - dwdq) < 1e-8, x


def test_kgrad_general_h():
    """For h != 1 the gradient factor scales as h^-5 (the h = 1 metric of :461-465 generalised)."""
    cfg = s.reference_config(8, 8, 8)
    cfg.h = 1.3
    cfg.stencil = 3
    k13 = scalar(cfg, "kgrad", 0.7, fp64=True)
    cfg.h = 1.0
    k1 = scalar(cfg, "kgrad", 0.7, fp64=True)
    assert abs(k13 / k1 - float(np.float32(1.3)) ** -5) < 1e-12


def test_sounding_hydrostatic():
    """Constant-N^2 sounding (asph_h100.cu:59-64, 136-148): dPi/dz = -g/(cp theta_ref),
    dp/dz = -rho g, rho_ref(0) = P0/(R TH0), Pi(0) = 1."""
    cfg = s.reference_config(8, 8, 8)
    th = lambda z: scalar(cfg, "theta_ref", z, fp64=True)
    ex = lambda z: scalar(cfg, "exner_ref", z, fp64=True)
    p = lambda z: scalar(cfg, "pres_ref", z, fp64=True)
    rho = lambda z: scalar(cfg, "rho_ref", z, fp64=True)
    assert ex(0.0) == 1.0 and th(0.0) == pytest.approx(300.0, abs=1e-12)
    assert rho(0.0) == pytest.approx(P0 / (R_D * 300.0), rel=1e-9)
    for z in (100.0, 2000.0, 7000.0):
        e = 1e-2
        dex = (ex(z + e) - ex(z - e)) / (2 * e)
        assert dex == pytest.approx(-GRAV / (CP * th(z)), rel=1e-6)
        dp = (p(z + e) - p(z - e)) / (2 * e)
        assert dp == pytest.approx(-rho(z) * GRAV, rel=1e-6)
    # FP32 instance within float rounding of FP64
    for z in (0.0, 1234.5, 7999.0):
        assert scalar(cfg, "rho_ref", z) == pytest.approx(rho(z), rel=2e-6)


def test_neutral_sounding_limit():
    """N^2 -> 0: theta_ref = TH0 and Pi = 1 - g z/(cp TH0)."""
    cfg = s.reference_config(8, 8, 8)
    cfg.n2 = 0.0
    z = 3000.0
    assert scalar(cfg, "theta_ref", z, fp64=True) == pytest.approx(300.0)
    assert scalar(cfg, "exner_ref", z, fp64=True) == pytest.approx(1 - GRAV * z / (CP * 300.0), rel=1e-12)
    cfg.n2 = 1e-9   # the exponential form approaches the limit
    assert scalar(cfg, "exner_ref", z, fp64=True) == pytest.approx(1 - GRAV * z / (CP * 300.0), rel=1e-6)


def test_magnus_and_qsat():
    """esat (Magnus, :155-158): 611.2 Pa at 0 C; qsat = 0.622 e/(p - 0.378 e) (:159-163);
    dqsdT is the T-derivative of qsat (:164-171)."""
    cfg = s.reference_config(8, 8, 8)
    assert scalar(cfg, "esat", 273.15, fp64=True) == pytest.approx(611.2, rel=1e-6)
    es30 = 611.2 * math.exp(17.67 * 30.0 / (30.0 + 243.5))
    assert scalar(cfg, "esat", 303.15, fp64=True) == pytest.approx(es30, rel=1e-6)
    T, p = 290.0, 90000.0
    es = scalar(cfg, "esat", T, fp64=True)
    assert scalar(cfg, "qsat", T, p, fp64=True) == pytest.approx(0.622 * es / (p - 0.378 * es), rel=1e-6)
    e = 1e-4
    d = (scalar(cfg, "qsat", T + e, p, fp64=True) - scalar(cfg, "qsat", T - e, p, fp64=True)) / (2 * e)
    assert scalar(cfg, "dqsdT", T, p, fp64=True) == pytest.approx(d, rel=1e-6)


def test_rh_profile_interpolation():
    cfg, ic, dt, _ = s.preset(3)
    for z, rh in ((0.0, 0.8), (500.0, 0.8), (1500.0, 1.01), (1999.0, 1.01), (2000.0, 0.8), (5000.0, 0.5),
                  (8000.0, 0.2), (9000.0, 0.2)):
        ex = scalar(cfg, "exner_ref", z, fp64=True)
        T = scalar(cfg, "theta_ref", z, fp64=True) * ex
        p = P0 * ex ** (CP / R_D)
        q = scalar(cfg, "qsat", T, p, fp64=True)
        assert scalar(cfg, "qv_ref", z, fp64=True) == pytest.approx(rh * q, rel=1e-5), z


def _one_particle_state(cfg, th, qv, qc):
    st = s.ic_generate(cfg, s.ICSpec(seed=1, jitter=0.0))
    st.thq[:, 0], st.thq[:, 1], st.thq[:, 2] = th, qv, qc
    return st


def test_saturation_adjustment_invariants():
    """Newton saturation adjustment (:370-385) conserves qv+qc and theta_l = theta - L qc/(cp Pi),
    and leaves qv ~ qs(T) when cloud remains."""
    cfg, ic, dt = small_config("moist", (8, 8, 6))
    cfg.flags &= ~(s._lib.F_AUTOCONV | s._lib.F_ACCRETION | s._lib.F_RAIN_EVAP)
    st = s.ic_generate(cfg, ic)
    rng = np.random.default_rng(3)
    st.thq[:, 1] *= rng.uniform(0.9, 1.3, cfg.n).astype(np.float32)     # sub- and supersaturated
    st.thq[:, 2] = rng.choice([0.0, 5e-4, 2e-3], cfg.n).astype(np.float32)
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.sort()
    pos0, vel0, thq0, rho0 = o.get_state_f64()
    o.density()
    o.eos_micro(dt)
    pos, vel, thq, rho = o.get_state_f64()
    _, prs, _ = o.get_aux_f64()
    z = pos[:, 2]
    ex = np.array([scalar(cfg, "exner_ref", zz, fp64=True) for zz in z])
    pref = P0 * ex ** (CP / R_D)
    pfl = np.maximum(pref + prs, 1.0)
    Pi = (pfl / P0) ** (R_D / CP)
    lv, cp = L_V, CP
    assert np.allclose(thq[:, 1] + thq[:, 2], thq0[:, 1] + thq0[:, 2], rtol=0, atol=1e-15)
    thl0 = thq0[:, 0] - lv * thq0[:, 2] / (cp * Pi)
    thl = thq[:, 0] - lv * thq[:, 2] / (cp * Pi)
    assert np.allclose(thl, thl0, rtol=0, atol=1e-9)
    cloudy = thq[:, 2] > 1e-6
    assert cloudy.sum() > 10
    qs = np.array([scalar(cfg, "qsat", T, p, fp64=True) for T, p in zip(thq[cloudy, 0] * Pi[cloudy], pfl[cloudy])])
    assert np.allclose(thq[cloudy, 1], qs, rtol=1e-6)


def test_autoconversion_and_rain_mass_balance():
    """Autoconversion (qc - qc0)(1 - exp(-k dt)) into the rain grid (:389-394); sedimentation
    (:577-599) conserves sum(rain) + sum(precip) exactly (fixed-point grid)."""
    cfg, ic, dt = small_config("moist", (8, 8, 8))
    cfg.flags &= ~(s._lib.F_ACCRETION | s._lib.F_RAIN_EVAP)
    st = s.ic_generate(cfg, ic)
    st.thq[:, 1] = 0.03          # supersaturated: the cloud survives the adjustment
    st.thq[:, 2] = 3e-3
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.sort()
    o.density()
    _, _, thq_before, _ = o.get_state_f64()
    o.eos_micro(dt)
    _, _, thq_after, _ = o.get_state_f64()
    rain, precip, clipped = o.get_rain_fx()
    # deposited rain (particle-mass units * 2^40) equals the qc removed beyond the adjustment
    dep = rain.sum() / 2.0**40
    assert dep > 0
    total = rain.sum() + precip.sum()
    for _ in range(30):
        o.rain(dt)
        r, p, c = o.get_rain_fx()
        assert c == 0
        assert r.sum() + p.sum() == total
    assert p.sum() > 0


def test_rain_fall_distance():
    """Upstream sedimentation moves frac = V dt / dz of a cell's rain down per step (:587-594)."""
    cfg, ic, dt = small_config("ref", (6, 6, 10))
    o = Oracle(cfg)
    o.set_state(s.ic_generate(cfg, ic))
    rain = np.zeros(cfg.n, np.int64)
    top = (cfg.nz - 1) * cfg.nx * cfg.ny
    rain[top] = 1 << 50
    o.set_rain_fx(rain)
    o.rain(dt)
    r, p, c = o.get_rain_fx()
    frac = min(np.float32(cfg.v_rain) * np.float32(dt) * np.float32(1.0 / cfg.dz), 1.0)
    moved = int((1 << 50) * float(np.float32(frac)))
    assert r[top] == (1 << 50) - moved and r[top - cfg.nx * cfg.ny] == moved


def test_kdk_free_particle_and_walls():
    """kick/drift (:535-561): x += v dt with a = 0; periodic wrap in x, y; reflection in z."""
    cfg, ic, dt = small_config("ref", (6, 6, 6))
    st = s.ic_generate(cfg, ic)
    st.vel[:, 0], st.vel[:, 1], st.vel[:, 2] = 10.0, -7.0, 0.5
    st.vel[0, :3] = [11500.0, 0.0, -600.0]     # wraps in x, reflects off the floor
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.kick_drift(dt)
    pos, vel, _, _ = o.get_state_f64()
    p0 = st.pos.astype(np.float64)
    exp_x = (p0[1:, 0] + 10.0 * dt) % (cfg.nx * cfg.dx)
    assert np.allclose(pos[1:, 0], exp_x, atol=1e-9)
    assert np.allclose(pos[1:, 1], (p0[1:, 1] - 7.0 * dt) % (cfg.ny * cfg.dy), atol=1e-9)
    x0 = p0[0, 0] + 11500.0 * dt
    assert pos[0, 0] == pytest.approx(x0 - cfg.nx * cfg.dx, abs=1e-9)
    z0 = p0[0, 2] - 600.0 * dt
    assert z0 < 0 and pos[0, 2] == pytest.approx(-z0, abs=1e-9) and vel[0, 2] == 600.0


def test_density_uniform_rho_old():
    """Lagged Shepard (:286-333): with a uniform old density the ratio returns it (up to the
    clamp [0.5, 2] rho_ref)."""
    cfg, ic, dt = small_config("dry", (8, 8, 8))
    st = s.ic_generate(cfg, ic)
    st.rho[:] = 1.0
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.sort()
    o.density()
    _, _, _, rho = o.get_state_f64()
    z = o.get_state_f64()[0][:, 2]
    rref = np.array([scalar(cfg, "rho_ref", zz, fp64=True) for zz in z])
    expect = np.clip(1.0, 0.5 * rref, 2.0 * rref)
    assert np.allclose(rho, expect, rtol=1e-13)


def test_momentum_rest_uniform_pressure_buoyancy():
    """Momentum (:415-511) at rest with uniform p': no pressure or viscous force, so acc is the
    buoyancy g[(thv - thvr)/thvr - qc] alone and dtheta/dt = 0."""
    cfg, ic, dt = small_config("moist", (8, 8, 8))
    st = s.ic_generate(cfg, ic)
    st.thq[:, 0] += 0.5
    st.thq[:, 2] = 1e-4
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.sort()
    o.set_prs(np.full(cfg.n, 37.0, np.float32))
    o.momentum()
    pos, vel, thq, rho = o.get_state_f64()
    acc, prs, dth = o.get_aux_f64()
    assert np.all(acc[:, :2] == 0.0)
    z = pos[:, 2]
    thr = np.array([scalar(cfg, "theta_ref", zz, fp64=True) for zz in z])
    qvr = np.array([scalar(cfg, "qv_ref", zz, fp64=True) for zz in z])
    thv = thq[:, 0] * (1 + float(np.float32(0.608)) * thq[:, 1])
    thvr = thr * (1 + float(np.float32(0.608)) * qvr)
    b = GRAV * ((thv - thvr) / thvr - thq[:, 2])
    assert np.allclose(acc[:, 2], b, rtol=1e-12, atol=1e-15)
    assert np.all(dth == 0.0)


def test_viscosity_gated_on_approach():
    """Monaghan viscosity acts only on approaching pairs (:480): two receding particles feel no
    viscous force, and viscous heating is zero."""
    cfg, ic, dt = small_config("ref", (8, 8, 8))
    st = s.ic_generate(cfg, ic)
    # radial expansion about the domain centre: every pair recedes
    c = np.array([cfg.nx * cfg.dx, cfg.ny * cfg.dy, cfg.nz * cfg.dz]) / 2
    st.vel[:, :3] = (1e-3 * (st.pos[:, :3] - c)).astype(np.float32)
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.sort()
    o.set_prs(np.zeros(cfg.n, np.float32))
    o.momentum()
    acc, prs, dth = o.get_aux_f64()
    assert np.all(dth == 0.0)
    assert np.all(acc[:, :2] == 0.0)
