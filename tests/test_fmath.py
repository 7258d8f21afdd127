"""Accuracy of the project's exp / log / pow (asph_fmath.h), which host and device share so that
their results agree bit for bit. Checked against float64 numpy on the ranges the physics uses."""
import numpy as np
import pytest

from oracle.oracle import load


def ulp_err(got, exact):
    got = np.asarray(got, np.float32)
    ref32 = np.asarray(exact, np.float64).astype(np.float32)
    sp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - np.asarray(exact, np.float64)) / sp


def vec(fn, *args):
    return np.array([fn(*a) for a in zip(*args)], np.float32)


def test_fexp_ulp():
    lib = load()
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-87, 88, 20000), rng.uniform(-2, 2, 20000),
                        np.array([0.0, 1.0, -1.0, 1e-8, -1e-8, 88.7, -87.3])]).astype(np.float32)
    got = vec(lib.oracle_fexp, x)
    err = ulp_err(got, np.exp(x.astype(np.float64)))
    assert err.max() <= 2.0, err.max()
    assert lib.oracle_fexp(0.0) == 1.0
    assert np.isinf(lib.oracle_fexp(100.0)) and lib.oracle_fexp(-200.0) == 0.0


def test_flog_ulp():
    lib = load()
    rng = np.random.default_rng(1)
    x = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), rng.uniform(0.5, 2.0, 20000),
                        np.array([1.0, 2.0, 0.5, 1e-30, 3e38])]).astype(np.float32)
    got = vec(lib.oracle_flog, x)
    err = np.abs(got.astype(np.float64) - np.log(x.astype(np.float64)))
    # absolute error relative to the spacing of the result (log near 1 -> near 0)
    ref = np.log(x.astype(np.float64)).astype(np.float32)
    tol = np.maximum(np.spacing(np.abs(ref)).astype(np.float64), 2e-8)
    assert (err / tol).max() <= 3.0, (err / tol).max()
    assert lib.oracle_flog(1.0) == 0.0
    assert np.isinf(lib.oracle_flog(0.0)) and np.isnan(lib.oracle_flog(-1.0))


@pytest.mark.parametrize("y,lo,hi", [(0.2859, 0.3, 1.1), (3.4977, 0.5, 1.05), (0.875, 1e-9, 1e-1),
                                      (0.525, 1e-9, 1e-2), (0.2046, 1e-9, 1e-2)])
def test_fpow_relative(y, lo, hi):
    """pow on the bases the step uses: the Exner function and p_ref (p/p0, Pi) and the Kessler
    rates (qr, rho*qr). Error bound ~ ulp * (1 + 3 |y ln x|) (the log error scales by y ln x)."""
    lib = load()
    rng = np.random.default_rng(2)
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), 10000)).astype(np.float32)
    got = vec(lambda a: lib.oracle_fpow(a, y), x)
    yy = np.float64(np.float32(y))
    exact = np.power(x.astype(np.float64), yy)
    rel = np.abs(got - exact) / exact
    bound = 1.5e-7 * (1.0 + 3.0 * np.abs(yy * np.log(x.astype(np.float64))))
    assert (rel / bound).max() <= 1.0, (rel / bound).max()
    assert lib.oracle_fpow(0.0, y) == 0.0 and lib.oracle_fpow(1.0, y) == 1.0
