"""The device adjoint (asph_adjoint_rain): against the host adjoint (same schedule, same pair sets)
along the FP32 trajectory and along the FP64 one, against central differences of the FP64 oracle,
through the torch.autograd binding, and for its side effects on the handle.

Tolerances. The device differentiates the FP32 model (the forward is the FP32 oracle's, bit for
bit) and the FP64 references differentiate the FP64 model, whose constants and trajectory differ
in the last float bits (the autoconversion factor 1 - exp(-k dt) by 2.4e-4 relative in float):
measured 5.5e-5 normwise per component on the 20x20x16 case. Along the FP32 trajectory, with the
stages the device keeps in float evaluated in float, the host adjoint differs from the device
only by summation order.
"""
import numpy as np
import pytest
import torch

import sph_cuda_amd as s
from sph_cuda_amd import _lib as L
from helpers import small_config
from oracle.oracle import adjoint_rain_f64, rain_objective_f64
from sph_cuda_amd.autograd import rain_objective

pytestmark = pytest.mark.gpu

SHAPE, NSTEPS, CK = (20, 20, 16), 8, 4


def _normwise(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _windy(cfg, st, cells, nsteps, dt):
    """Uniform wind that carries every particle `cells` cells along x (half along y) over the window."""
    vel = st.vel.copy()
    vel[:, 0] += np.float32(cells * cfg.dx / (nsteps * dt))
    vel[:, 1] += np.float32(0.5 * cells * cfg.dy / (nsteps * dt))
    return s.State(st.pos, vel, st.thq, st.rho)


@pytest.fixture(scope="module")
def case():
    cfg, ic, dt = small_config("adj", SHAPE)
    st = s.ic_generate(cfg, ic)
    print(f"[case] adjoint {SHAPE} steps {NSTEPS} checkpoint {CK}", flush=True)
    with s.Sim(cfg) as sim:
        sim.set_state(st)
        r = sim.adjoint_rain(NSTEPS, dt, CK)
    return cfg, dt, st, r, adjoint_rain_f64(cfg, st, NSTEPS, dt, CK)


def test_matches_host_adjoint(case):
    cfg, dt, st, r, (J64, g64, gq64) = case
    g = r["grad_thq"].astype(np.float64)
    rel = [_normwise(g[:, c], g64[:, c]) for c in range(3)]
    print(f"[case] normwise theta/qv/qc {rel}, J {r['J']:.9e} vs {J64:.9e}, "
          f"qc0 {r['dJ_dqc0']:.6e} vs {gq64:.6e}", flush=True)
    assert r["segments"] == (NSTEPS + CK - 1) // CK
    assert np.all(np.isfinite(g))
    assert max(rel) <= 5e-4
    assert abs(r["J"] - J64) <= 5e-3 * J64
    assert abs(r["dJ_dqc0"] - gq64) <= 5e-4 * abs(gq64)
    assert np.array_equal(r["grad_theta"], r["grad_thq"][:, 0])


# the reverse stages the device evaluates in float: both pair stages, the momentum epilogue, the
# density ratio and the state adjoints (the EOS and microphysics stage runs in double)
DEVICE_FLOAT_STAGES = 1 | 2 | 8 | 16 | 32


def test_matches_host_adjoint_on_fp32_trajectory(case):
    cfg, dt, st, r, _ = case
    J32, g32, gq32 = adjoint_rain_f64(cfg, st, NSTEPS, dt, CK, trajectory="f32",
                                      float_stages=DEVICE_FLOAT_STAGES)
    g = r["grad_thq"].astype(np.float64)
    rel = [_normwise(g[:, c], g32[:, c]) for c in range(3)]
    print(f"[case] FP32 trajectory: normwise theta/qv/qc {rel}, J {r['J']:.12e} vs {J32:.12e}, "
          f"qc0 {r['dJ_dqc0']:.9e} vs {gq32:.9e}", flush=True)
    assert r["J"] == J32   # the same forward, bit for bit
    assert max(rel) <= 2e-5
    assert abs(r["dJ_dqc0"] - gq32) <= 2e-5 * abs(gq32)


# (compression of the 20 x 20 x 16 adjoint case towards the bubble centre, which keeps the cloud
# at its height): the reverse pair stages' halo exceeds the LDS capacity and is streamed in slabs
# of whole z-layers (x3 in z: ~1300 particles per layer), or of halo rows where one layer alone
# exceeds it (x16 in z: all 6400 particles in one layer)
DENSE = [(1.0, 1.0, 3.0), (1.0, 1.0, 16.0)]


@pytest.mark.parametrize("a", DENSE)
def test_dense_halo_slabs(a):
    """Compressed states whose reverse pair stages stream their halos through LDS in slabs: the
    device against the host adjoint along the FP32 trajectory (the same forward bit for bit)."""
    cfg, ic, dt = small_config("adj", SHAPE)
    st = s.ic_generate(cfg, ic)
    c = np.array([ic.bubble_x, ic.bubble_y, ic.bubble_z], np.float32)
    st.pos[:, :3] = (c + (st.pos[:, :3] - c) / np.asarray(a, np.float32)).astype(np.float32)
    nsteps, ck = 4, 2
    print(f"[case] adjoint {SHAPE} compressed {a} steps {nsteps} checkpoint {ck}", flush=True)
    with s.Sim(cfg) as sim:
        sim.set_state(st)
        r = sim.adjoint_rain(nsteps, dt, ck)
    J32, g32, gq32 = adjoint_rain_f64(cfg, st, nsteps, dt, ck, trajectory="f32",
                                      float_stages=DEVICE_FLOAT_STAGES)
    g = r["grad_thq"].astype(np.float64)
    rel = [_normwise(g[:, k], g32[:, k]) for k in range(3)]
    print(f"[case] normwise theta/qv/qc {rel}, J {r['J']:.12e} vs {J32:.12e}", flush=True)
    assert J32 > 0
    assert r["J"] == J32
    assert np.all(np.isfinite(g))
    assert max(rel) <= 2e-5
    assert abs(r["dJ_dqc0"] - gq32) <= 2e-5 * abs(gq32)


def test_matches_central_differences(case):
    cfg, dt, st, r, _ = case
    g = r["grad_theta"].astype(np.float64)
    rng = np.random.default_rng(4000)
    base = st.thq.astype(np.float64)
    eps = 1e-3
    fds, ads = [], []
    for k in range(3):
        d = rng.standard_normal(cfg.n)
        tp, tm = base.copy(), base.copy()
        tp[:, 0] += eps * d
        tm[:, 0] -= eps * d
        fd = (rain_objective_f64(cfg, st.pos, st.vel, tp, st.rho, NSTEPS, dt, CK)
              - rain_objective_f64(cfg, st.pos, st.vel, tm, st.rho, NSTEPS, dt, CK)) / (2 * eps)
        fds.append(fd)
        ads.append(float(g @ d))
        print(f"[case] dir {k}: FD {fd:.6e} adjoint {ads[-1]:.6e}", flush=True)
    fds, ads = np.array(fds), np.array(ads)
    assert _normwise(ads, fds) <= 1e-3


def test_stale_lists_beyond_one_cell():
    """One segment with a wind that moves every particle 2.5 cells from its listed cell: the forward
    leaves the staged bricks for the global cell table and the reverse walk widens its margin with
    the recorded drift bound. The pair sets must stay those of the host adjoint."""
    cfg, ic, dt = small_config("adj", SHAPE)
    n = 8
    st = _windy(cfg, s.ic_generate(cfg, ic), 2.5, n, dt)
    print(f"[case] windy adjoint {SHAPE} steps {n}", flush=True)
    with s.Sim(cfg) as sim:
        sim.set_state(st)
        r = sim.adjoint_rain(n, dt, n)
    J64, g64, gq64 = adjoint_rain_f64(cfg, st, n, dt, n)
    g = r["grad_thq"].astype(np.float64)
    rel = [_normwise(g[:, c], g64[:, c]) for c in range(3)]
    print(f"[case] windy normwise theta/qv/qc {rel}, qc0 {r['dJ_dqc0']:.6e} vs {gq64:.6e}", flush=True)
    assert max(rel) <= 1e-3
    assert abs(r["dJ_dqc0"] - gq64) <= 1e-3 * abs(gq64)


def test_handle_state_is_restored(case):
    cfg, dt, st, _, _ = case
    with s.Sim(cfg) as sim, s.Sim(cfg) as ref:
        sim.set_state(st)
        ref.set_state(st)
        sim.adjoint_rain(3, dt, 2)
        a, b = sim.get_state(), ref.get_state()
        for f in ("pos", "vel", "thq", "rho"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        for h in (sim, ref):
            h.sort()
            h.force(dt)
            h.step(dt, 0.0, do_sort=False)
        a, b = sim.get_state(), ref.get_state()
        for f in ("pos", "vel", "thq", "rho"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f


def test_autograd_binding(case):
    cfg, dt, st, r, _ = case
    theta0 = torch.tensor(st.thq[:, 0], dtype=torch.float32, requires_grad=True)
    qc0 = torch.tensor(float(cfg.qc_auto), dtype=torch.float64, requires_grad=True)
    J = rain_objective(theta0, qc0, st, cfg, NSTEPS, dt, CK)
    assert J.dtype == torch.float64 and abs(float(J) - r["J"]) <= 1e-12 * r["J"]
    (3.0 * J).backward()
    assert theta0.grad is not None and qc0.grad is not None
    # a second run of the adjoint: its FP64 atomics (rain grid, qc0) may sum in another order
    g3 = 3.0 * r["grad_theta"].astype(np.float64)
    np.testing.assert_allclose(theta0.grad.numpy(), g3, rtol=1e-4, atol=1e-6 * np.max(np.abs(g3)))
    assert float(qc0.grad) == pytest.approx(3.0 * r["dJ_dqc0"], rel=1e-6)


def test_autograd_device_inputs_and_forward_only(case):
    """theta0 and qc0 on the GPU: J and both gradients come back on the GPU with the values of the
    host path; without gradients the forward window alone gives the same J bit for bit."""
    cfg, dt, st, r, _ = case
    dev = torch.device("cuda:0")
    theta0 = torch.tensor(st.thq[:, 0], dtype=torch.float32, device=dev, requires_grad=True)
    qc0 = torch.tensor(float(cfg.qc_auto), dtype=torch.float64, device=dev, requires_grad=True)
    J = rain_objective(theta0, qc0, st, cfg, NSTEPS, dt, CK)
    assert J.device.type == "cuda" and float(J.detach()) == r["J"]
    J.backward()
    assert theta0.grad.device.type == "cuda" and qc0.grad.device.type == "cuda"
    g = r["grad_theta"].astype(np.float64)
    np.testing.assert_allclose(theta0.grad.cpu().numpy(), g, rtol=1e-4, atol=1e-6 * np.max(np.abs(g)))
    assert float(qc0.grad) == pytest.approx(r["dJ_dqc0"], rel=1e-6)
    with torch.no_grad():
        J0 = rain_objective(theta0.detach(), qc0.detach(), st, cfg, NSTEPS, dt, CK)
    J1 = rain_objective(torch.tensor(st.thq[:, 0]), float(cfg.qc_auto), st, cfg, NSTEPS, dt, CK)
    assert float(J0) == float(J1) == r["J"]


def test_rejects_what_it_cannot_differentiate(case):
    cfg8, ic8, dt = small_config("adj", (8, 8, 8))   # narrower than 2 * stencil + 3
    with s.Sim(cfg8) as sim:
        sim.set_state(s.ic_generate(cfg8, ic8))
        with pytest.raises(L.AsphError, match="nx, ny"):
            sim.adjoint_rain(2, dt, 2)
    cfg, dt, st, _, _ = case
    with s.Sim(cfg) as sim:
        sim.set_state(st)
        with pytest.raises(L.AsphError, match="nsteps"):
            sim.adjoint_rain(0, dt, 2)
