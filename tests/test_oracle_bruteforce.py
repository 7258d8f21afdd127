"""The FP64 oracle's neighbour sums against an independent float64 restatement of the
reference formulas evaluated over ALL particle pairs (O(N^2)) on small lattices.

Right after a sort every particle sits in the cell it is listed in, and the stencil covers the
support, so the reference's stencil loop (asph_h100.cu:306-327, 444-495) visits exactly the
pairs with r^2 < 4h^2. The restatement below writes the reference expressions literally
(dw3(q)/q, (pj - pi)/(rho_i rho_j), mu = v.r/(rp^2 + eta^2), visc = -alpha cs H mu/(rbar^2 rho_j));
the oracle's canonical forms must agree with it to double rounding."""
import numpy as np
import pytest

import sph_cuda_amd as s
from helpers import small_config
from oracle.oracle import Oracle

F = lambda v: float(np.float32(v))
GRAV, R_D, CP, P0 = F(9.80665), F(287.05), F(1004.0), F(1.0e5)
SIGMA = F(0.31830988618)


def sounding(cfg, z):
    th0, n2 = F(cfg.th0), F(cfg.n2)
    th = th0 * np.exp(n2 * z / GRAV)
    A = GRAV * GRAV / (CP * th0 * n2)
    ex = 1.0 - A * (1.0 - np.exp(-n2 * z / GRAV))
    p = P0 * ex ** (CP / R_D)
    return th, ex, p, p / (R_D * th * ex)


def qv_ref(cfg, z):
    if cfg.qv_mode == s._lib.QV_REF_EXP:
        q = F(cfg.qv0) * np.exp(-z / F(cfg.qv_hscale))
        return q * np.exp(-np.maximum(z - F(cfg.qv_ztop), 0.0) / F(cfg.qv_htop))
    assert cfg.qv_mode == s._lib.QV_DRY
    return np.zeros_like(z)


def w3(q):
    return np.where(q >= 2, 0.0, np.where(q < 1, SIGMA * (1 - 1.5 * q * q + 0.75 * q ** 3),
                                          0.25 * SIGMA * (2 - q) ** 3))


def dw3(q):
    return np.where(q >= 2, 0.0, np.where(q < 1, SIGMA * (-3 * q + 2.25 * q * q), -0.75 * SIGMA * (2 - q) ** 2))


def pair_geometry(cfg, pos):
    d = np.array([F(cfg.dx), F(cfg.dy), F(cfg.dz)])
    r = pos[:, None, :3] - pos[None, :, :3]            # r_ij = x_i - x_j
    if cfg.flags & s._lib.F_MIN_IMAGE:
        for a, n in ((0, cfg.nx), (1, cfg.ny)):
            L = n * d[a]
            r[..., a] = np.where(r[..., a] > 0.5 * L, r[..., a] - L, r[..., a])
            r[..., a] = np.where(r[..., a] < -0.5 * L, r[..., a] + L, r[..., a])
    xi = r / d
    r2 = (xi ** 2).sum(-1)
    return r, xi, r2, d


def brute_density(cfg, pos, rho_old):
    h = F(cfg.h)
    _, _, r2, _ = pair_geometry(cfg, pos)
    W = np.where(r2 < 4 * h * h, w3(np.sqrt(r2) / h), 0.0)
    S1 = W.sum(1)
    S2 = (W / np.maximum(rho_old, 1e-3)[None, :]).sum(1)
    _, _, _, rref = sounding(cfg, pos[:, 2])
    rr = np.where(S2 > 1e-20, S1 / S2, rref)
    return np.clip(rr, 0.5 * rref, 2.0 * rref)


def brute_momentum(cfg, pos, vel, thq, rho, prs):
    h = F(cfg.h)
    r, xi, r2, d = pair_geometry(cfg, pos)
    ok = (r2 < 4 * h * h) & (r2 >= 1e-12)
    q = np.sqrt(r2) / h
    with np.errstate(divide="ignore", invalid="ignore"):
        k = np.where(ok, dw3(q) / q, 0.0) * h ** -5
    g = k[..., None] * xi / d
    rhoc = np.maximum(rho, 1e-3)
    dp = (prs[None, :] - prs[:, None]) / (rhoc[:, None] * rhoc[None, :])
    vd = vel[:, None, :3] - vel[None, :, :3]
    vdotr = (vd * r).sum(-1)
    hav = F(cfg.h_av)
    eta2 = (0.01 * hav) ** 2
    rp = np.sqrt((r ** 2).sum(-1))
    mu = vdotr / (rp * rp + eta2)
    rbar = 0.5 * (rhoc[:, None] + rhoc[None, :])
    visc = np.where(vdotr < 0, -F(cfg.alpha_av) * F(cfg.cs) * hav * mu / (rbar * rbar * rhoc[None, :]), 0.0)
    visc = np.where(ok, visc, 0.0)
    br = np.where(ok, dp + visc, 0.0)
    vcell = F(cfg.dx) * F(cfg.dy) * F(cfg.dz)
    rho0 = P0 / (R_D * F(cfg.th0))
    mm = rho0 * vcell / vcell
    a = -(br[..., None] * g).sum(1) * mm
    av = -(visc[..., None] * g).sum(1) * mm
    th, ex, p, _ = sounding(cfg, pos[:, 2])
    thv = thq[:, 0] * (1 + F(0.608) * thq[:, 1])
    thvr = th * (1 + F(0.608) * qv_ref(cfg, pos[:, 2]))
    a[:, 2] += GRAV * ((thv - thvr) / thvr - thq[:, 2])
    Pi = (np.maximum(p + prs, 1.0) / P0) ** (R_D / CP)
    dth = -(vel[:, :3] * av).sum(1) / (CP * Pi)
    return a, np.clip(dth, -F(cfg.dthdt_cap), F(cfg.dthdt_cap))


CASES = [("ref", (7, 7, 6)), ("ref", (6, 8, 7)), ("dry", (8, 8, 8)), ("dry", (9, 7, 8))]


@pytest.mark.parametrize("kind,n", CASES)
@pytest.mark.parametrize("min_image", [True, False])
def test_density_and_momentum_vs_all_pairs(kind, n, min_image):
    cfg, ic, dt = small_config(kind, n)
    if min_image:
        cfg.flags |= s._lib.F_MIN_IMAGE
    else:
        cfg.flags &= ~s._lib.F_MIN_IMAGE
    st = s.ic_generate(cfg, ic)
    rng = np.random.default_rng(sum(n))
    st.vel[:, :3] = rng.normal(0, 2.0, (cfg.n, 3)).astype(np.float32)
    st.rho *= (1 + 0.01 * rng.standard_normal(cfg.n)).astype(np.float32)
    st.thq[:, 0] += rng.normal(0, 0.5, cfg.n).astype(np.float32)
    o = Oracle(cfg, fp64=True)
    o.set_state(st)
    o.sort()
    pos, vel, thq, rho_old = o.get_state_f64()
    o.density()
    _, _, _, rho = o.get_state_f64()
    assert np.allclose(rho, brute_density(cfg, pos, rho_old), rtol=1e-12, atol=0)
    prs = rng.normal(0, 50.0, cfg.n).astype(np.float32)
    o.set_prs(prs)
    o.momentum()
    acc, _, dth = o.get_aux_f64()
    a_ref, dth_ref = brute_momentum(cfg, pos, vel, thq, rho, prs.astype(np.float64))
    scale = np.abs(a_ref[:, :3]).max()
    assert np.abs(acc[:, :3] - a_ref).max() <= 1e-11 * scale
    assert np.allclose(dth, dth_ref, rtol=1e-9, atol=1e-15)
