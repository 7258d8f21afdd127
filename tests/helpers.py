"""Shared helpers of the parity tests: small configurations, comparison metrics."""
from __future__ import annotations

import numpy as np

import sph_cuda_amd as s
from sph_cuda_amd import _lib as L


def small_config(kind: str, n=(12, 12, 10)):
    """Test configurations on small lattices (nx, ny >= 2*stencil+1).

    kind: "ref"   - the reference's constants, flags and h = 1 (stencil 2);
          "dry"   - the 32^3 dry-bubble physics on a smaller lattice (h = 1.3, stencil 3);
          "moist" - the 128^3 moist physics (RH profile, all warm-rain processes, 2:1);
          "flux"  - "ref" with the surface flux on and a moist reference profile.
    """
    nx, ny, nz = n
    if kind == "ref":
        cfg = s.reference_config(nx, ny, nz)
        ic = L.ICSpec()
        ic.reference_init = 1
        ic.ref_noise = 0.1
        return cfg, ic, 0.5
    if kind == "dry":
        cfg, ic, dt, _ = s.preset(1)
    elif kind == "moist":
        cfg, ic, dt, _ = s.preset(2)
    elif kind == "adj":
        cfg, ic, dt, _ = s.preset(4)
    else:
        raise ValueError(kind)
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    lx, ly, lz = nx * cfg.dx, ny * cfg.dy, nz * cfg.dz
    ic.bubble_x, ic.bubble_y = 0.5 * lx, 0.5 * ly
    ic.bubble_z = min(ic.bubble_z, 0.4 * lz)
    ic.bubble_r = min(ic.bubble_r, 0.3 * lz)
    return cfg, ic, dt


def relnorm(a, b) -> float:
    """Normwise relative difference ||a-b|| / ||b|| (0 when both are 0)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    if nb == 0:
        return float(d)
    return float(d / nb)


def state_close(a: s.State, b: s.State, tol: float):
    out = {}
    for name in ("pos", "vel", "thq", "rho"):
        x, y = getattr(a, name), getattr(b, name)
        if name in ("pos", "vel", "thq"):
            for c in range(3):
                out[f"{name}.{c}"] = relnorm(x[:, c], y[:, c])
        else:
            out[name] = relnorm(x, y)
    bad = {k: v for k, v in out.items() if v > tol}
    return bad, out

