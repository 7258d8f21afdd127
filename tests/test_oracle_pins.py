"""Independent pins of the oracle functions that the device alone used to check (no GPU).

Each test evaluates a published formula in its own units or arrangement (not the oracle's), or a
numpy brute force, and compares the FP64 oracle with it:
  - Kessler rain evaporation: Klemp & Wilhelmson (1978), eq. 2.14, rho in g cm^-3 and p in mb;
  - Kessler accretion: dqc/dt = -k2 qc qr^0.875 (Kessler 1969; KW78 eq. 2.13), k2 = 2.2 s^-1,
    integrated backward-Euler in qc;
  - the bulk surface flux of asph_h100.cu:606-626 with the smoothstep skin temperature of
    asph_h100.cu:989-992 and the constants of asph_h100.cu:100-105;
  - the implicit Rayleigh sponge of asph_h100.cu:518-530 (constants :89-90);
  - the coarse block means and column diagnostics of asph_h100.cu:680-723.
"""
import numpy as np

import sph_cuda_amd as s
from helpers import small_config
from oracle.oracle import Oracle, scalar

L = s._lib
f32 = np.float32
# the reference's float constants (asph_h100.cu:66-70) as exact doubles
GRAV, R_D, CP, P0, L_V = (float(f32(v)) for v in (9.80665, 287.05, 1004.0, 1.0e5, 2.501e6))
FX64 = 2.0 ** 56            # rain grid unit of the FP64 oracle: particle mass * 2^-56


def _exner(cfg, z):
    return np.array([scalar(cfg, "exner_ref", float(zz), fp64=True) for zz in z])


def _pres_ref(cfg, z):
    return P0 * _exner(cfg, z) ** (CP / R_D)


def _qsat(cfg, T, p):
    return np.array([scalar(cfg, "qsat", float(a), float(b), fp64=True) for a, b in zip(T, p)])


def _cells(cfg, pos):
    d = np.array([cfg.dx, cfg.dy, cfg.dz], np.float64)
    n = np.array([cfg.nx, cfg.ny, cfg.nz])
    ijk = np.clip(np.floor(pos[:, :3] / d).astype(np.int64), 0, n - 1)
    return ijk, (ijk[:, 2] * cfg.ny + ijk[:, 1]) * cfg.nx + ijk[:, 0]


def _micro_case(flags_on, flags_off, qv_scale, qc, rain_qr, dt, seed=11, n=(8, 8, 6)):
    """Runs density + EOS/microphysics of the FP64 oracle on the moist small lattice with the
    given qv scaling, cloud water and rain mixing ratio per cell; returns the state before and
    after, the thermodynamic inputs the microphysics sees, and qr per particle."""
    cfg, ic, _ = small_config("moist", n)
    cfg.flags |= flags_on
    cfg.flags &= ~flags_off
    st = s.ic_generate(cfg, ic)
    rng = np.random.default_rng(seed)
    st.thq[:, 1] = (st.thq[:, 1] * qv_scale(rng, cfg.n)).astype(f32)
    st.thq[:, 2] = qc
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.sort()
    pos0, _, thq0, _ = o.get_state_f64()
    # rain per cell as a mixing ratio at the reference density, converted to the grid's unit
    rho0 = P0 / (R_D * float(f32(cfg.th0)))
    qr_cell = rain_qr(rng, cfg.n)
    rain_fx = np.rint(qr_cell / rho0 * FX64).astype(np.int64)      # kg/m^3 of rain / rho0
    o.set_rain_fx(rain_fx)
    o.density()
    o.eos_micro(dt)
    pos, _, thq, rho = o.get_state_f64()
    _, prs, _ = o.get_aux_f64()
    z = pos[:, 2]
    pfl = np.maximum(_pres_ref(cfg, z) + prs, 1.0)
    Pi = (pfl / P0) ** (R_D / CP)
    _, cid = _cells(cfg, pos)
    r = np.maximum(rho, 1e-3)
    qr = rain_fx[cid].astype(np.float64) / FX64 * rho0 / r       # kg rain / kg air
    return cfg, thq0, thq, r, pfl, Pi, qr


def _kw78_evaporation_rate(rho_si, qr, qv, qs, p_pa):
    """KW78 eq. 2.14 in the paper's units: rho [g cm^-3], p [mb]; returns E [1/s]."""
    rho_g = rho_si * 1e-3
    p_mb = p_pa / 100.0
    rq = rho_g * qr
    C = 1.6 + 124.9 * rq ** 0.2046                    # ventilation factor
    return (1.0 / rho_g) * (1.0 - qv / qs) * C * rq ** 0.525 / (5.4e5 + 2.55e6 / (p_mb * qs))


EVAP_ONLY = (L.F_RAIN_EVAP, L.F_AUTOCONV | L.F_ACCRETION)


def test_kessler_evaporation_kw78_paper_units():
    """Rain evaporation in subsaturated air follows KW78 eq. 2.14 (evaluated with rho in g/cm^3
    and p in mb, 2.55e6), removes rain and cools by L dqe / (cp Pi)."""
    dt = 0.25
    cfg, thq0, thq, r, pfl, Pi, qr = _micro_case(
        *EVAP_ONLY, qv_scale=lambda g, n: g.uniform(0.2, 0.95, n), qc=0.0,
        rain_qr=lambda g, n: g.choice([0.0, 1e-4, 5e-4, 2e-3], n) * g.uniform(0.5, 1.5, n), dt=dt)
    qv0, th0 = thq0[:, 1], thq0[:, 0]
    qs = _qsat(cfg, th0 * Pi, pfl)
    assert np.all(qv0 < qs)                                  # no saturation adjustment here
    E = np.where(qr > 0, _kw78_evaporation_rate(r, qr, qv0, qs, pfl), 0.0)
    dqe = np.minimum(np.minimum(E * dt, qr), qs - qv0)
    live = dqe > 0
    assert live.sum() > 0.5 * cfg.n * 0.75 * 0.5              # most rainy particles evaporate
    np.testing.assert_allclose(thq[:, 1] - qv0, dqe, rtol=2e-6, atol=1e-18)
    # latent cooling, in the sign and size of L dqe / (cp Pi)
    dth = thq[:, 0] - th0
    assert np.all(dth[live] < 0)
    np.testing.assert_allclose(dth, -L_V * dqe / (CP * Pi), rtol=2e-6, atol=1e-12)
    assert np.all(thq[:, 2] == 0.0)
    # the comparison separates the paper's constant from a units slip (2.55e6 with p in Pa)
    slip = np.minimum(np.minimum(np.where(qr > 0, _kw78_evaporation_rate(r, qr, qv0, qs, 100 * pfl), 0) * dt,
                                 qr), qs - qv0)
    assert np.median(np.abs(slip[live] / dqe[live] - 1)) > 0.1


def test_kessler_evaporation_caps():
    """With a long step the evaporated water is capped by the saturation deficit qs - qv (much
    rain) and by the rain present qr (little rain); both caps bind somewhere."""
    dt = 2.0e4
    cfg, thq0, thq, r, pfl, Pi, qr = _micro_case(
        *EVAP_ONLY, qv_scale=lambda g, n: g.uniform(0.9, 1.2, n), qc=0.0,
        rain_qr=lambda g, n: g.choice([1e-9, 1e-2], n), dt=dt, seed=5)
    qv0, th0 = thq0[:, 1], thq0[:, 0]
    qs = _qsat(cfg, th0 * Pi, pfl)
    E = _kw78_evaporation_rate(r, qr, qv0, qs, pfl)
    got = thq[:, 1] - qv0
    cap_def = (qs - qv0 < qr) & (qs - qv0 < E * dt)
    cap_qr = (qr < qs - qv0) & (qr < E * dt)
    assert cap_def.sum() > 20 and cap_qr.sum() > 20
    np.testing.assert_allclose(got[cap_def], (qs - qv0)[cap_def], rtol=1e-12, atol=1e-17)
    np.testing.assert_allclose(got[cap_qr], qr[cap_qr], rtol=1e-12, atol=1e-17)  # ulps of qv
    assert np.all(thq[:, 1] <= qs * (1 + 1e-12))             # never beyond saturation


def test_kessler_evaporation_zero_in_saturated_air():
    """No rain evaporates where the air is saturated after the adjustment (cloudy particles)."""
    dt = 0.25
    args = dict(qv_scale=lambda g, n: g.uniform(1.3, 1.6, n), qc=1e-3,
                rain_qr=lambda g, n: np.full(n, 2e-3), dt=dt)
    _, _, with_evap, *_ = _micro_case(*EVAP_ONLY, **args)
    _, _, no_evap, *_ = _micro_case(0, L.F_RAIN_EVAP | L.F_AUTOCONV | L.F_ACCRETION, **args)
    assert np.all(no_evap[:, 2] > 0)                          # cloud remains everywhere
    np.testing.assert_allclose(with_evap[:, 1], no_evap[:, 1], rtol=1e-12)


def test_kessler_accretion_rate_and_implicit_update():
    """Accretion removes cloud water at k2 qc qr^0.875 (k2 = 2.2 /s) in the small-dt limit, by the
    backward-Euler update qc / (1 + k2 dt qr^0.875) at any dt (never below zero), and changes
    neither theta nor qv."""
    assert f32(small_config("moist")[0].k_accr) == f32(2.2)
    acc_only = (L.F_ACCRETION, L.F_AUTOCONV | L.F_RAIN_EVAP)
    none = (0, L.F_ACCRETION | L.F_AUTOCONV | L.F_RAIN_EVAP)
    for dt, rtol_rate in ((1e-3, 1e-4), (30.0, None), (1e5, None)):
        args = dict(qv_scale=lambda g, n: g.uniform(1.2, 1.5, n), qc=2e-3,
                    rain_qr=lambda g, n: g.choice([0.0, 2e-4, 1e-3, 4e-3], n), dt=dt)
        _, _, a, _, _, _, qr = _micro_case(*acc_only, **args)
        _, _, b, _, _, _, _ = _micro_case(*none, **args)
        qc_pre = b[:, 2]                                       # cloud after the adjustment
        assert np.all(qc_pre > 0)
        np.testing.assert_array_equal(a[:, 1], b[:, 1])
        np.testing.assert_array_equal(a[:, 0], b[:, 0])
        expect = qc_pre / (1.0 + 2.2 * dt * qr ** 0.875)
        np.testing.assert_allclose(a[:, 2], expect, rtol=2e-6)
        assert np.all(a[:, 2] > 0) and np.all(a[:, 2] <= qc_pre)
        if rtol_rate:
            rain = qr > 0
            rate = (qc_pre - a[:, 2])[rain] / dt
            np.testing.assert_allclose(rate, 2.2 * qc_pre[rain] * qr[rain] ** 0.875, rtol=rtol_rate)
            assert np.all(a[~rain, 2] == qc_pre[~rain])


def _flux_case(simt, dt, th_set=None, qv_set=0.005, lift=()):
    """Reference configuration (surface flux on, autoconversion off): after a sort, particles in
    `lift` move up by 1.2 dz; then the flux + density + EOS stage at time simt."""
    cfg, ic, _ = small_config("ref", (8, 8, 6))
    cfg.flags &= ~L.F_AUTOCONV
    st = s.ic_generate(cfg, ic)
    st.thq[:, 1] = qv_set
    st.thq[:, 2] = 0.0
    if th_set is not None:
        st.thq[:, 0] = th_set
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.sort()
    pos, vel, thq, rho = o.get_state_f64()
    pos[list(lift), 2] += 1.2 * cfg.dz
    o.set_state_f64(pos, vel, thq, rho)
    o.density_eos(dt, simt)
    return cfg, pos, thq, o.get_state_f64()[2]


def test_surface_flux_bulk_formula():
    """dtheta = C_H |V| (Ts(t) - Ta) dt / dz and dqv = beta C_E |V| (qs(Ts, p0) - qv) dt / dz in
    the lowest layer, with Ts = TH0 + DTS_SKIN * smoothstep(t / RAMP); nothing elsewhere."""
    C_H, C_E, BETA, V, DTS, RAMP = 1.2e-3, 1.2e-3, 0.20, 7.0, 3.15, 3600.0   # asph_h100.cu:100-105
    cfg0 = small_config("ref")[0]
    assert [f32(v) for v in (cfg0.c_h, cfg0.c_e, cfg0.beta_moist, cfg0.v_bg, cfg0.dts_skin, cfg0.ramp)] == \
        [f32(v) for v in (C_H, C_E, BETA, V, DTS, RAMP)]
    dt = 0.5
    for simt, sw in ((0.0, 0.0), (900.0, 0.15625), (1800.0, 0.5), (3600.0, 1.0), (7200.0, 1.0)):
        cfg, pos, th_before, th_after = _flux_case(simt, dt, lift=(0, 1, 2))
        Ts = float(f32(cfg.th0)) + DTS * sw
        qs_sfc = scalar(cfg, "qsat", Ts, float(_pres_ref(cfg, [0.0])[0]), fp64=True)
        z = pos[:, 2]
        Pi = (np.maximum(_pres_ref(cfg, z), 1.0) / P0) ** (R_D / CP)
        Ta = th_before[:, 0] * Pi
        lifted = np.zeros(cfg.n, bool)
        lifted[[0, 1, 2]] = True                               # listed at iz = 0, now above dz
        assert np.all(z[lifted] >= cfg.dz)
        bottom = (z < cfg.dz) & ~lifted
        assert bottom.sum() == cfg.nx * cfg.ny - 3
        dth = C_H * V * (Ts - Ta) * dt / cfg.dz
        dqv = BETA * C_E * V * (qs_sfc - th_before[:, 1]) * dt / cfg.dz
        got_th = th_after[:, 0] - th_before[:, 0]
        got_qv = th_after[:, 1] - th_before[:, 1]
        np.testing.assert_allclose(got_th[bottom], dth[bottom], rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(got_qv[bottom], dqv[bottom], rtol=1e-6, atol=1e-15)
        assert np.all(got_th[~bottom] == 0) and np.all(got_qv[~bottom] == 0)
        if sw > 0:
            assert np.all(got_th[bottom] > 0)                 # a warm skin heats the layer


def test_surface_flux_clamps():
    """An overshooting flux is clamped to theta in [100, THETA_CAP] and qv in [0, QV_CAP]."""
    _, pos, b, a = _flux_case(7200.0, 1e6, th_set=290.0, qv_set=0.0)
    bottom = pos[:, 2] < 100.0
    assert np.all(a[bottom, 0] == 500.0) and np.all(a[bottom, 1] == float(f32(0.06)))
    _, pos, b, a = _flux_case(7200.0, 1e6, th_set=400.0, qv_set=0.05)
    bottom = pos[:, 2] < 100.0
    assert np.all(a[bottom, 0] == 100.0) and np.all(a[bottom, 1] == 0.0)


def test_sponge_implicit_factor():
    """Above zs = (1 - 0.15) Lz the vertical velocity is multiplied by 1 / (1 + 0.05 t^2 dt), with
    t = (z - zs) / (0.15 Lz); below, velocities are untouched."""
    SIGMA, FRAC = 0.05, 0.15                                  # asph_h100.cu:89-90
    cfg0 = small_config("ref")[0]
    assert f32(cfg0.sponge_sigma) == f32(SIGMA) and f32(cfg0.sponge_frac) == f32(FRAC)
    dt = 0.5
    outs = []
    for sponge in (True, False):
        cfg, ic, _ = small_config("ref", (8, 8, 10))
        if not sponge:
            cfg.flags &= ~L.F_SPONGE
        st = s.ic_generate(cfg, ic)
        rng = np.random.default_rng(2)
        st.vel[:, :3] = rng.normal(0, 3.0, (cfg.n, 3)).astype(f32)
        o = Oracle(cfg, fp64=True)
        o.set_state(st)
        o.sort()
        o.force(dt)
        o.momentum_kdk(dt)
        outs.append(o.get_state_f64())
    (pos, vel_on, _, _), (_, vel_off, _, _) = outs
    Lz = cfg.nz * cfg.dz
    zs = (1 - FRAC) * Lz
    z = pos[:, 2]
    t = np.clip((z - zs) / (FRAC * Lz), 0, None)
    factor = np.where(z >= zs, 1.0 / (1.0 + SIGMA * t * t * dt), 1.0)
    top = z >= zs
    assert top.sum() > 50 and (~top).sum() > 50
    np.testing.assert_allclose(vel_on[:, 2], vel_off[:, 2] * factor, rtol=1e-7, atol=1e-12)
    assert np.all(vel_on[~top, 2] == vel_off[~top, 2])
    np.testing.assert_array_equal(vel_on[:, :2], vel_off[:, :2])


def _diag_state(n=(10, 9, 7)):
    cfg, ic, _ = small_config("moist", n)
    st = s.ic_generate(cfg, ic)
    rng = np.random.default_rng(9)
    st.vel[:, :3] = rng.normal(0, 2.0, (cfg.n, 3)).astype(f32)
    st.thq[:, 2] = rng.choice([0.0, 1e-4, 3e-3], cfg.n).astype(f32)
    st.pos[:, 2] *= f32(0.55)              # the top layers empty: empty blocks and columns' tops
    o = Oracle(cfg)
    o.set_state(st)
    o.sort()
    g = o.get_state()
    pos, vel, thq = (x.astype(np.float64) for x in (g.pos, g.vel, g.thq))
    return cfg, o, pos, vel, thq


def test_coarse_dump_block_means_bruteforce():
    """Block means of [theta, qv, qc, u, v, w] over sx x sy x sz cell blocks (partial blocks at the
    high edges, empty blocks 0), block index cxx + cx (cyy + cy czz), against numpy."""
    cfg, o, pos, vel, thq = _diag_state()
    for sx, sy, sz in ((4, 4, 3), (1, 1, 1), (3, 9, 2)):
        got = o.coarse_dump(sx, sy, sz)
        cx, cy, cz = -(-cfg.nx // sx), -(-cfg.ny // sy), -(-cfg.nz // sz)
        ijk, _ = _cells(cfg, pos)
        b = ((ijk[:, 2] // sz) * cy + ijk[:, 1] // sy) * cx + ijk[:, 0] // sx
        fields = np.stack([thq[:, 0], thq[:, 1], thq[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]], 1)
        sums = np.zeros((cx * cy * cz, 6))
        np.add.at(sums, b, fields)
        cnt = np.bincount(b, minlength=cx * cy * cz).astype(np.float64)
        want = np.where(cnt[:, None] > 0, sums / np.maximum(cnt, 1)[:, None], 0.0).reshape(cz, cy, cx, 6)
        assert (cnt == 0).sum() > 0
        scale = np.abs(fields).max(0)
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6 * scale.max())
        assert np.all(got.reshape(-1, 6)[cnt == 0] == 0.0)


def test_column_diag_bruteforce():
    """Per column: mean theta, mean qv, max qc, max |w| over every particle listed in the column."""
    cfg, o, pos, vel, thq = _diag_state()
    got = o.column_diag()
    ijk, _ = _cells(cfg, pos)
    col = ijk[:, 1] * cfg.nx + ijk[:, 0]
    ncol = cfg.nx * cfg.ny
    cnt = np.bincount(col, minlength=ncol)
    assert np.all(cnt > 0)
    mth = np.bincount(col, thq[:, 0], ncol) / cnt
    mqv = np.bincount(col, thq[:, 1], ncol) / cnt
    qcmax = np.zeros(ncol)
    np.maximum.at(qcmax, col, thq[:, 2])
    wmax = np.zeros(ncol)
    np.maximum.at(wmax, col, np.abs(vel[:, 2]))
    np.testing.assert_allclose(got[:, 0], mth, rtol=2e-6)
    np.testing.assert_allclose(got[:, 1], mqv, rtol=2e-6)
    np.testing.assert_array_equal(got[:, 2], qcmax.astype(f32))
    np.testing.assert_array_equal(got[:, 3], wmax.astype(f32))
