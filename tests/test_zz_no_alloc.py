"""asph_coarse_dump_noalloc (fixed per-GPU staging array, no device allocation) against
asph_coarse_dump on the same handles, through the C ABI: bit for bit, on two calls in a row with
the second smaller, and a dump beyond the fixed array rejected with an error."""
import ctypes as C

import numpy as np
import pytest

import sph_cuda_amd as s
from helpers import small_config
from sph_cuda_amd import _lib as L

pytestmark = pytest.mark.gpu


def _handle(n, steps):
    cfg, ic, dt = small_config("moist", n)
    sim = s.Sim(cfg)
    sim.set_state(s.ic_generate(cfg, ic))
    if steps:
        rc, _ = sim.run(max_steps=steps, dt=dt, fixed_dt=True)
        assert rc == 0
    sim.sort()
    return cfg, sim


def _dump(lib, fn, sim, cfg, stride, fill):
    sx, sy, sz = stride
    cx, cy, cz = -(-cfg.nx // sx), -(-cfg.ny // sy), -(-cfg.nz // sz)
    out = np.full(cx * cy * cz * 6, fill, np.float32)
    rc = getattr(lib, fn)(sim.h, sx, sy, sz, out.ctypes.data_as(C.POINTER(C.c_float)))
    return rc, out


def test_noalloc_coarse_dump_matches_allocating_dump():
    lib = L.load()
    cfg1, big = _handle((24, 24, 16), 3)
    cfg2, small = _handle((12, 12, 10), 2)
    cases = [(big, cfg1, (1, 1, 1)), (small, cfg2, (2, 2, 3)), (big, cfg1, (5, 4, 3))]
    sizes = []
    for sim, cfg, stride in cases:
        rc_a, want = _dump(lib, "asph_coarse_dump", sim, cfg, stride, np.nan)
        rc_b, got = _dump(lib, "asph_coarse_dump_noalloc", sim, cfg, stride, -7.0)
        assert rc_a == 0 and rc_b == 0, L.last_error()
        assert np.array_equal(got, want), (stride, np.abs(got - want).max())
        assert np.all(np.isfinite(got)) and np.any(got != 0)
        sizes.append(got.size)
    assert sizes[1] < sizes[0]


def test_noalloc_coarse_dump_rejects_oversize():
    lib = L.load()
    cfg, sim = _handle((64, 64, 48), 0)
    n = 64 * 64 * 48 * 6
    assert n > L.NOALLOC_MAX_FLOATS
    out = np.zeros(n, np.float32)
    rc = lib.asph_coarse_dump_noalloc(sim.h, 1, 1, 1, out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == L.ERR and "fixed buffer" in L.last_error()
    # the same handle still dumps at strides that fit
    rc, got = _dump(lib, "asph_coarse_dump_noalloc", sim, cfg, (2, 2, 2), 0.0)
    rc2, want = _dump(lib, "asph_coarse_dump", sim, cfg, (2, 2, 2), 0.0)
    assert rc == 0 and rc2 == 0 and np.array_equal(got, want)
