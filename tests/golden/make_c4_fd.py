"""Writes c4_fd.npz: the finite-difference and FP64-adjoint reference of the 64^3 rain adjoint.

Workload: the 64^3 adjoint preset (262,144 particles, moist, initial cloud in the bubble), a
60-step window of dt = 0.25 s with the neighbour structure rebuilt every 10 steps, objective J =
rain water in the domain at the end of the window. For 64 directions d_k (k = 0..63, standard
normal per particle from numpy's default_rng(4000 + k)) it records:
  fd_1e3[k], fd_1e5[k] - central differences (J(theta0 + eps d) - J(theta0 - eps d)) / (2 eps) of
                         the FP64 oracle, eps = 1e-3 K and 1e-5 K;
  gd_host[k]           - g . d_k with g = dJ/dtheta0 from the FP64 host adjoint (oracle/);
plus J and dJ/dqc0 of the FP64 host adjoint and a digest of the initial state.

Runs on the CPU with the oracle library (about 80 minutes on 8 cores):
    python tests/golden/make_c4_fd.py
Finished directions are kept in c4_fd.progress.json beside the output, so an interrupted run
resumes where it stopped; the file is removed when c4_fd.npz is written.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import sph_cuda_amd as s  # noqa: E402
from oracle.oracle import adjoint_rain_f64, rain_objective_f64  # noqa: E402

NDIR, SEED0, EPS = 64, 4000, (1e-3, 1e-5)


def state_digest(st) -> str:
    h = hashlib.sha256()
    for a in (st.pos, st.vel, st.thq, st.rho):
        h.update(np.ascontiguousarray(a, np.float32).tobytes())
    return h.hexdigest()


def directions(n):
    for k in range(NDIR):
        yield k, np.random.default_rng(SEED0 + k).standard_normal(n)


def main():
    cfg, ic, dt, nsteps = s.preset(4)
    ck = 10
    st = s.ic_generate(cfg, ic)
    digest = state_digest(st)
    prog_path = os.path.join(HERE, "c4_fd.progress.json")
    prog = json.load(open(prog_path)) if os.path.exists(prog_path) else {}
    if prog.get("state_sha256") != digest:
        prog = {"state_sha256": digest, "dirs": {}}
    t0 = time.time()
    if "J" not in prog:
        J, g, gq = adjoint_rain_f64(cfg, st, nsteps, dt, ck)
        prog.update(J=J, dJ_dqc0=gq)
        prog["gd_host"] = [float(g[:, 0] @ d) for _, d in directions(cfg.n)]
        json.dump(prog, open(prog_path, "w"))
        print(f"host adjoint: J {J:.12e} dJ/dqc0 {gq:.9e} ({time.time() - t0:.0f} s)", flush=True)
    base = st.thq.astype(np.float64)
    for k, d in directions(cfg.n):
        if str(k) in prog["dirs"]:
            continue
        fd = []
        for e in EPS:
            tp, tm = base.copy(), base.copy()
            tp[:, 0] += e * d
            tm[:, 0] -= e * d
            jp = rain_objective_f64(cfg, st.pos, st.vel, tp, st.rho, nsteps, dt, ck)
            jm = rain_objective_f64(cfg, st.pos, st.vel, tm, st.rho, nsteps, dt, ck)
            fd.append((jp - jm) / (2 * e))
        prog["dirs"][str(k)] = fd
        json.dump(prog, open(prog_path, "w"))
        print(f"dir {k}: g.d {prog['gd_host'][k]:.9e} fd(1e-3) {fd[0]:.9e} fd(1e-5) {fd[1]:.9e} "
              f"({time.time() - t0:.0f} s)", flush=True)
    fd = np.array([prog["dirs"][str(k)] for k in range(NDIR)])
    np.savez(os.path.join(HERE, "c4_fd.npz"), J=prog["J"], dJ_dqc0=prog["dJ_dqc0"],
             gd_host=np.array(prog["gd_host"]), fd_1e3=fd[:, 0], fd_1e5=fd[:, 1], seed0=SEED0,
             nsteps=nsteps, dt=dt, checkpoint=ck, state_sha256=np.array(digest))
    os.remove(prog_path)


if __name__ == "__main__":
    main()
