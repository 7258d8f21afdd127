"""Reader and writer of the reference's coarse dump files `asph_t%06.0fs.bin`.

Format (asph_h100.cu:743-747, 1106-1120; documented in tools/make_viz.py:4-10): a 68-byte
little-endian header `<2I10i5f` = magic 0x41535048, version 3, nx, ny, nz, cx, cy, cz, sx, sy,
sz, nfields (6), t_sim, dt, lx, ly, lz; then cx*cy*cz*6 float32 values [theta, qv, ql, u, v, w]
per coarse cell, cell index cxx + cx*(cyy + cy*czz); an empty cell is all zeros.
"""
from __future__ import annotations

import struct

import numpy as np

MAGIC = 0x41535048
VERSION = 3
HEADER = struct.Struct("<2I10i5f")
FIELDS = ("theta", "qv", "ql", "u", "v", "w")


def read_dump(path: str) -> dict:
    with open(path, "rb") as f:
        d = f.read()
    h = HEADER.unpack(d[: HEADER.size])
    magic, version, nx, ny, nz, cx, cy, cz, sx, sy, sz, nf, t, dt, lx, ly, lz = h
    if magic != MAGIC:
        raise ValueError(f"bad magic in {path}")
    vals = np.frombuffer(d, dtype="<f4", count=cx * cy * cz * nf, offset=HEADER.size)
    data = vals.reshape(cz, cy, cx, nf)
    return dict(version=version, nx=nx, ny=ny, nz=nz, cx=cx, cy=cy, cz=cz, sx=sx, sy=sy, sz=sz,
                nfields=nf, t=t, dt=dt, lx=lx, ly=ly, lz=lz, data=data,
                fields={name: data[..., i] for i, name in enumerate(FIELDS[:nf])})


def write_dump(path: str, nx, ny, nz, sx, sy, sz, t_sim, dt, lx, ly, lz, data: np.ndarray) -> None:
    cx, cy, cz = -(-nx // sx), -(-ny // sy), -(-nz // sz)
    arr = np.ascontiguousarray(data, dtype="<f4").reshape(cz, cy, cx, 6)
    with open(path, "wb") as f:
        f.write(HEADER.pack(MAGIC, VERSION, nx, ny, nz, cx, cy, cz, sx, sy, sz, 6,
                            t_sim, dt, lx, ly, lz))
        f.write(arr.tobytes())


def dump_name(outdir: str, simt: float) -> str:
    return "%s/asph_t%06.0fs.bin" % (outdir, simt)


def viz_frames(paths) -> dict:
    """The JSON payload the reference's visualiser builds from a list of dumps
    (tools/make_viz.py:39-56): per-frame theta/qv/ql/w rounded to 2/5/5/3 decimals."""
    frames = [read_dump(p) for p in sorted(paths)]
    t0 = frames[0]
    th = t0["fields"]["theta"]

    def r(a, nd):
        return [[[round(float(v), nd) for v in row] for row in sl] for sl in a]

    return dict(
        Lx=t0["lx"], Ly=t0["ly"], Lz=t0["lz"], cx=t0["cx"], cy=t0["cy"], cz=t0["cz"],
        dx=t0["lx"] / t0["nx"], dy=t0["ly"] / t0["ny"], dz=t0["lz"] / t0["nz"],
        theta_ref=round(float(th[0, 0, 0]), 2),
        times=[round(float(fr["t"]), 1) for fr in frames],
        theta=[r(fr["fields"]["theta"], 2) for fr in frames],
        qv=[r(fr["fields"]["qv"], 5) for fr in frames],
        ql=[r(fr["fields"]["ql"], 5) for fr in frames],
        w=[r(fr["fields"]["w"], 3) for fr in frames],
    )
