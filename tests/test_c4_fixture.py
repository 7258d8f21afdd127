"""The 64^3 finite-difference reference (tests/golden/c4_fd.npz) on the CPU: it belongs to the
adjoint workload's initial state, and the FP64 host adjoint (oracle/asph_adjoint_ref.cpp) agrees
with central differences of the FP64 oracle along its 64 directions, which pins the host adjoint at
that size before the GPU test compares the device with both."""
import hashlib
import os

import numpy as np

import sph_cuda_amd as s

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c4_fd.npz")


def _normwise(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_fixture_matches_the_workload():
    ref = np.load(FIX)
    cfg, ic, dt, nsteps = s.preset(4)
    assert int(ref["nsteps"]) == nsteps == 60 and int(ref["checkpoint"]) == 10
    assert np.float32(ref["dt"]) == np.float32(dt)
    st = s.ic_generate(cfg, ic)
    h = hashlib.sha256()
    for a in (st.pos, st.vel, st.thq, st.rho):
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    assert h.hexdigest() == str(ref["state_sha256"])
    assert len(ref["gd_host"]) == len(ref["fd_1e3"]) == len(ref["fd_1e5"]) == 64
    assert float(ref["J"]) > 0 and float(ref["dJ_dqc0"]) != 0


def test_host_adjoint_agrees_with_central_differences():
    ref = np.load(FIX)
    # at eps = 1e-5 K the differences resolve the derivative of the FP64 model
    assert _normwise(ref["gd_host"], ref["fd_1e5"]) <= 5e-4
    # the 1e-5 K and 1e-3 K differences agree with each other only to the piecewise smoothness
    # of the objective over 60 steps (switches of saturation, autoconversion and clamps)
    assert _normwise(ref["fd_1e3"], ref["fd_1e5"]) <= 2e-2
