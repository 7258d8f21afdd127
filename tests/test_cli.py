"""The command-line driver (sph_cuda_amd/asph_cli, built from sph_cuda_amd/csrc/asph_cli.cpp): the
reference program's flags (asph_h100.cu:754-788), its report lines (:1121-1175) and dumps."""
import os
import re
import subprocess

import numpy as np
import pytest

from sph_cuda_amd.dumpfmt import read_dump

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph_cuda_amd", "asph_cli")
FLAGS = ("--nx", "--ny", "--nz", "--dt", "--cfl", "--cfl-interval", "--max-trip", "--strict-cfl",
         "--dump-every", "--sim-hours", "--dump-stride", "--outdir", "--noise", "--validate", "--quick")


def _need_cli():
    if not os.path.exists(CLI):
        pytest.skip("asph_cli is not built (make -C sph_cuda_amd/csrc)")


def test_help_lists_the_reference_flags():
    _need_cli()
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    for f in FLAGS:
        assert f in r.stdout, f


def test_bad_arguments_exit_1_before_touching_a_device():
    _need_cli()
    r = subprocess.run([CLI, "--no-such-flag"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unknown flag" in r.stderr
    r = subprocess.run([CLI, "--nx"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    r = subprocess.run([CLI, "--nx", "2000", "--ny", "8", "--nz", "8"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ERROR" in r.stderr     # the 1024-cells-per-axis key limit


DUMP_LINE = re.compile(r"^\d+\s+\+\d+\s*s\s+[\d.]+\s+[\d.]+\s+CFL [\d.]+ \| cloud\s+[\d.]+% ql_max [\d.]+ g/kg \| "
                       r"w_max\s+[\d.]+ m/s \| precip [\d.]+ mm$")


@pytest.mark.gpu
def test_run_prints_the_reference_reports_and_writes_dumps(tmp_path):
    _need_cli()
    out = str(tmp_path / "dumps")
    cmd = [CLI, "--nx", "16", "--ny", "16", "--nz", "16", "--sim-hours", "0.0025", "--dump-every", "3",
           "--dump-stride", "4", "4", "--validate", "--report", "5", "--outdir", out]
    print("[case] " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    dumps = [l for l in lines if DUMP_LINE.match(l)]
    assert len(dumps) >= 2
    assert sum(l.startswith("  AUDIT") for l in lines) == len(dumps)
    assert sum(l.startswith("  DRIFT") for l in lines) == len(dumps)
    assert any(l.startswith("  [step ") for l in lines)
    assert "=== DONE" in r.stdout and "parcel-steps :" in r.stdout
    files = sorted(f for f in os.listdir(out) if f.endswith(".bin"))
    assert len(files) == len(dumps)
    d = read_dump(os.path.join(out, files[-1]))
    assert (d["nx"], d["ny"], d["nz"], d["sx"], d["sz"], d["nfields"]) == (16, 16, 16, 4, 4, 6)
    assert np.all(np.isfinite(d["data"]))


@pytest.mark.gpu
def test_run_hook_events_and_stop():
    """asph_run_hooked: dump and report events with the loop state; a true return stops the run."""
    import sph_cuda_amd as s
    from helpers import small_config

    cfg, ic, dt = small_config("ref", (12, 12, 10))
    events = []

    def hook(ev):
        events.append((ev.kind, ev.step, round(ev.simt, 6), ev.dt))
        return ev.kind == s._lib.EV_REPORT and ev.step >= 8

    with s.Sim(cfg) as g:
        g.set_state(s.ic_generate(cfg, ic))
        rc, st = g.run(max_steps=50, dt=dt, fixed_dt=True, dump_every=2 * dt, hook=hook, report_every=4)
    print(f"[case] events {events}", flush=True)
    assert rc == 0 and st.steps == 8
    reports = [e for e in events if e[0] == s._lib.EV_REPORT]
    dumps = [e for e in events if e[0] == s._lib.EV_DUMP]
    assert [e[1] for e in reports] == [4, 8]
    assert [e[1] for e in dumps] == [2, 4, 6, 8]
    assert all(abs(e[2] - e[1] * dt) < 1e-6 for e in events)
