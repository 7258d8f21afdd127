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
