"""The C ABI (include/asph.h) without a GPU: the library loads, exports every declared entry
point, the binding's struct layouts match, and the host-only entry points (configuration,
presets, initial conditions) behave as specified."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import sph_cuda_amd as s
from sph_cuda_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "asph.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(asph_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert len(names) >= 40
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_struct_layouts(lib):
    assert lib.asph_sizeof_config() == C.sizeof(L.Config)
    assert lib.asph_sizeof_ic_spec() == C.sizeof(L.ICSpec)
    assert lib.asph_sizeof_run_args() == C.sizeof(L.RunArgs)
    assert lib.asph_sizeof_run_stats() == C.sizeof(L.RunStats)


def test_reference_config_constants():
    """asph_h100.cu:48-120 and Config defaults :728-742."""
    c = s.reference_config(1000, 1000, 150)
    assert (c.dx, c.dy, c.dz, c.h, c.stencil) == (1000.0, 1000.0, 100.0, 1.0, 2)
    assert (c.th0, c.cs) == (300.0, 80.0) and c.n2 == np.float32(6e-5)
    assert c.h_av == 200.0 and abs(c.alpha_av - 0.2) < 1e-7
    assert c.qc_auto == np.float32(1e-3) and c.v_rain == 9.0 and c.theta_cap == 500.0
    assert abs(c.cfl - 0.4) < 1e-7 and c.max_trip == 20 and c.strict_cfl == 0
    assert c.flags & L.F_SURFACE_FLUX and c.flags & L.F_SPONGE
    assert not c.flags & L.F_GATHER_ACC and not c.flags & L.F_MIN_IMAGE


def test_config_check_errors(lib):
    c = s.reference_config(1025, 8, 8)   # asph_h100.cu:785-788
    assert lib.asph_config_check(C.byref(c)) == L.ERR
    assert b"1024" in lib.asph_last_error()
    c = s.reference_config(4, 8, 8)      # stencil 2 needs nx >= 5
    assert lib.asph_config_check(C.byref(c)) == L.ERR
    c = s.reference_config(8, 8, 8)
    c.h, c.stencil = 1.3, 2              # stencil does not cover the support
    assert lib.asph_config_check(C.byref(c)) == L.ERR
    c = s.reference_config(8, 8, 8)
    assert lib.asph_config_check(C.byref(c)) == L.OK
    with pytest.raises(s.AsphError):
        s.check_config(s.reference_config(2000, 8, 8))


@pytest.mark.parametrize("which,n,dt,steps", [(1, 32, 0.5, 120), (2, 128, 0.25, 600), (3, 256, 0.2, 3000),
                                             (4, 64, 0.25, 60), (5, 32, 0.25, 240)])
def test_presets(which, n, dt, steps):
    cfg, ic, pdt, psteps = s.preset(which)
    assert cfg.nx == n and pdt == pytest.approx(dt) and psteps == steps
    s.check_config(cfg)
    assert cfg.h == pytest.approx(1.3) and cfg.flags & L.F_MIN_IMAGE and cfg.flags & L.F_GATHER_ACC


def test_ic_deterministic_and_jittered():
    cfg, ic, dt = s.preset(1)[0], s.preset(1)[1], 0.5
    cfg.nx = cfg.ny = cfg.nz = 16
    a = s.ic_generate(cfg, ic)
    b = s.ic_generate(cfg, ic)
    for f in ("pos", "vel", "thq", "rho"):
        assert np.array_equal(getattr(a, f), getattr(b, f))
    ic.seed = 99
    c = s.ic_generate(cfg, ic)
    assert not np.array_equal(a.pos, c.pos)
    # lattice + U(+-0.1 dx) jitter
    n = cfg.nx
    idx = np.arange(cfg.n)
    ix, iy, iz = idx % n, (idx // n) % n, idx // (n * n)
    jit = a.pos[:, 0] / cfg.dx - (ix + 0.5)
    assert np.abs(jit).max() <= 0.1 + 1e-5 and np.abs(jit).max() > 0.09
    assert np.abs(a.pos[:, 2] / cfg.dz - (iz + 0.5)).max() <= 0.1 + 1e-5
    assert np.all(a.vel == 0) and np.all(a.thq[:, 1:] == 0)   # dry: qv = qc = 0


def test_ic_bubble_and_noise():
    cfg, ic, dt, _ = s.preset(3)
    cfg.nx = cfg.ny = 32
    cfg.nz = 32
    st = s.ic_generate(cfg, ic)
    from oracle.oracle import scalar

    th_ref = np.array([scalar(cfg, "theta_ref", z) for z in st.pos[:, 2]])
    noise = st.thq[:, 0] - th_ref
    assert np.sqrt(np.mean(noise.astype(np.float64) ** 2)) == pytest.approx(0.2, rel=1e-3)  # RMS renormalised
    # correlated: neighbouring lattice sites in x are strongly correlated (L = 500 m = 4 cells)
    nz = noise.reshape(cfg.nz, cfg.ny, cfg.nx)
    r = np.corrcoef(nz[:, :, :-1].ravel(), nz[:, :, 1:].ravel())[0, 1]
    assert r > 0.9
    cfg, ic, dt, _ = s.preset(1)
    st = s.ic_generate(cfg, ic)
    th_ref = np.array([scalar(cfg, "theta_ref", z) for z in st.pos[:, 2]])
    d = st.thq[:, 0] - th_ref
    r = np.linalg.norm(st.pos[:, :3] - np.array([4000.0, 4000.0, 2000.0]), axis=1)
    assert d.max() == pytest.approx(2.0, abs=0.05) and np.all(d[r > 2000] == 0)
    inside = r < 1000
    assert np.allclose(d[inside], 2.0 * np.cos(np.pi * r[inside] / 4000.0) ** 2, atol=1e-4)


def test_reference_init_matches_kernel_init():
    """reference_init reproduces kernel_init (asph_h100.cu:210-227): cell centres, hash noise."""
    cfg = s.reference_config(8, 8, 8)
    ic = L.ICSpec()
    ic.reference_init, ic.ref_noise = 1, 0.1
    st = s.ic_generate(cfg, ic)
    i = np.arange(cfg.n)
    ix, iy, iz = i % 8, (i // 8) % 8, i // 64
    assert np.array_equal(st.pos[:, 0], ((ix + 0.5) * 1000.0).astype(np.float32))
    assert np.array_equal(st.pos[:, 2], ((iz + 0.5) * 100.0).astype(np.float32))

    def hash_noise(a, b, c):
        h = (np.uint32(a) * np.uint32(374761393) + np.uint32(b) * np.uint32(668265263)
             + np.uint32(c) * np.uint32(2246822519)) & np.uint32(0xffffffff)
        h = np.uint32((h ^ (h >> np.uint32(13))) * np.uint32(1274126177))
        h ^= h >> np.uint32(16)
        return np.float32(h & np.uint32(0xffffff)) * np.float32(1.0 / 16777215.0) - np.float32(0.5)

    from oracle.oracle import scalar

    with np.errstate(over="ignore"):
        for k in (0, 77, 511):
            th = np.float32(scalar(cfg, "theta_ref", float(st.pos[k, 2]))) + np.float32(0.1) * hash_noise(ix[k], iy[k], iz[k])
            assert st.thq[k, 0] == pytest.approx(float(th), abs=2e-5)


def test_sim_requires_gpu_or_raises():
    """Creating a device handle without a GPU fails cleanly (no exit, an error message)."""
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    cfg = s.reference_config(8, 8, 8)
    with pytest.raises(s.AsphError):
        s.Sim(cfg)
