"""The asph_t*.bin dump format (asph_h100.cu:743-747, 1106-1120; tools/make_viz.py:4-37) against
dumps the reference shipped (tests/golden/, written by tests/golden/make_fixtures.py)."""
import json
import os

import numpy as np
import pytest

from sph_cuda_amd.dumpfmt import HEADER, MAGIC, dump_name, read_dump, write_dump

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_header_is_68_bytes():
    assert HEADER.size == 68


@pytest.mark.parametrize("f,grid,coarse,stride,size", [
    ("asph_t000060s.bin", (128, 128, 64), (13, 13, 22), (10, 10, 3), 89300),
    ("asph_t000300s.bin", (160, 160, 80), (16, 16, 27), (10, 10, 3), 165956),
    ("dumps_new_t300.bin", (128, 128, 64), (13, 13, 22), (10, 10, 3), 89300),
])
def test_read_headers(f, grid, coarse, stride, size):
    path = os.path.join(G, f)
    d = read_dump(path)
    assert os.path.getsize(path) == size == 68 + 4 * 6 * np.prod(coarse)
    assert (d["nx"], d["ny"], d["nz"]) == grid
    assert (d["cx"], d["cy"], d["cz"]) == coarse
    assert (d["sx"], d["sy"], d["sz"]) == stride
    assert d["version"] == 3 and d["nfields"] == 6
    # cx = ceil(nx / sx) (asph_h100.cu:853-855); L = n * D with the reference's 1000/1000/100 m
    assert d["cx"] == -(-d["nx"] // d["sx"]) and d["cz"] == -(-d["nz"] // d["sz"])
    assert (d["lx"], d["lz"]) == (grid[0] * 1000.0, grid[2] * 100.0)


def test_dump_values_known():
    d = read_dump(os.path.join(G, "asph_t000060s.bin"))
    assert d["t"] == pytest.approx(60.32578659057617) and d["dt"] == pytest.approx(0.33255794644355774)
    n = read_dump(os.path.join(G, "dumps_new_t300.bin"))
    th = n["fields"]["theta"]
    assert np.all(th > 0)                          # every coarse cell populated
    assert th.min() == pytest.approx(300.75, abs=0.01) and th.max() == pytest.approx(336.45, abs=0.01)
    w = n["fields"]["w"]
    assert w.min() == pytest.approx(-41.8, abs=0.05) and w.max() == pytest.approx(66.9, abs=0.05)


def test_visualiser_conversion_matches_sph_data():
    """The rounding/transposition of tools/make_viz.py:45-56 reproduces sph_data.json exactly."""
    exp = json.load(open(os.path.join(G, "viz_frames.json")))
    for f, fr in exp["frames"].items():
        d = read_dump(os.path.join(G, f))
        assert round(float(d["t"]), 1) == fr["t"]
        for name, nd in (("theta", 2), ("qv", 5), ("ql", 5), ("w", 3)):
            got = [[[round(float(v), nd) for v in row] for row in sl] for sl in d["fields"][name]]
            assert got == fr[name], (f, name)
    d0 = read_dump(os.path.join(G, "asph_t000060s.bin"))
    assert exp["top"]["theta_ref"] == round(float(d0["fields"]["theta"][0, 0, 0]), 2)
    assert exp["top"]["dx"] == d0["lx"] / d0["nx"]


def test_roundtrip_byte_identical(tmp_path):
    for f in ("asph_t000060s.bin", "asph_t000300s.bin"):
        src = os.path.join(G, f)
        d = read_dump(src)
        out = str(tmp_path / f)
        write_dump(out, d["nx"], d["ny"], d["nz"], d["sx"], d["sy"], d["sz"], d["t"], d["dt"],
                   d["lx"], d["ly"], d["lz"], d["data"])
        assert open(out, "rb").read() == open(src, "rb").read()


def test_file_name():
    """"%s/asph_t%06.0fs.bin" of simt (asph_h100.cu:1107)."""
    assert dump_name("d", 60.32578659057617) == "d/asph_t000060s.bin"
    assert dump_name("d", 2400.4) == "d/asph_t002400s.bin"
    assert dump_name("d", 299.5) == "d/asph_t000300s.bin"
    assert MAGIC == 0x41535048
