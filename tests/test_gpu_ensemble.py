"""GPU: ensembles (several domains in one handle, one launch per stage) and hipGraph replay.

An ensemble member must evolve exactly as the same domain run alone with the same sort
schedule, and graph replay must reproduce eager stepping bit for bit."""
import numpy as np
import pytest

import sph_cuda_amd as s
from helpers import small_config

pytestmark = pytest.mark.gpu

# members of different shapes; the 12 x 12 x 10 one is narrower than a brick halo (thread-per-
# particle path), the others take the LDS brick path
SHAPES = [(16, 16, 12), (20, 18, 12), (12, 12, 10), (24, 24, 14)]


def _member(k, shape):
    cfg, ic, dt = small_config("moist", shape)
    ic.seed = 100 + k
    ic.white_noise = 0.1
    return cfg, ic, dt


def _concat(states):
    return s.State(*[np.concatenate([getattr(x, f) for x in states]) for f in ("pos", "vel", "thq", "rho")])


def _schedule(n):
    return [k % 2 == 0 for k in range(n)]


def _eager(sim, dt, n):
    for k, srt in enumerate(_schedule(n)):
        sim.step(dt, k * dt, do_sort=srt)


def test_ensemble_members_match_single_runs():
    mem = [_member(k, sh) for k, sh in enumerate(SHAPES)]
    dt = mem[0][2]
    states = [s.ic_generate(c, ic) for c, ic, _ in mem]
    ens = s.Sim([c for c, _, _ in mem])
    assert ens.members == len(SHAPES)
    ens.set_state(_concat(states))
    ens.sort()
    ens.force(dt)
    _eager(ens, dt, 6)
    e_state = ens.get_state()
    e_rain, e_prec, _ = ens.get_rain_fx()
    col = 0
    for m, (c, ic, _) in enumerate(mem):
        print(f"[case] member {m} shape {SHAPES[m]}", flush=True)
        one = s.Sim(c)
        one.set_state(states[m])
        one.sort()
        one.force(dt)
        _eager(one, dt, 6)
        o_state = one.get_state()
        a, n = ens.member_range(m)
        for f in ("pos", "vel", "thq", "rho"):
            assert np.array_equal(getattr(e_state, f)[a:a + n], getattr(o_state, f)), (m, f)
        r, p, _ = one.get_rain_fx()
        assert np.array_equal(e_rain[a:a + n], r)
        assert np.array_equal(e_prec[col:col + c.nx * c.ny], p)
        col += c.nx * c.ny
        assert np.array_equal(ens.member_max_vel(m), one.max_vel())


@pytest.mark.parametrize("ensemble", [False, True])
def test_graph_replay_matches_eager(ensemble):
    mem = [_member(k, sh) for k, sh in enumerate(SHAPES if ensemble else SHAPES[:1])]
    dt = mem[0][2]
    st = _concat([s.ic_generate(c, ic) for c, ic, _ in mem])
    cfgs = [c for c, _, _ in mem] if ensemble else mem[0][0]
    eager, graph = s.Sim(cfgs), s.Sim(cfgs)
    for x in (eager, graph):
        x.set_state(st)
        x.sort()
        x.force(dt)
    _eager(eager, dt, 8)
    stats = graph.run_graph(8, dt, sort_every=2)
    assert stats.steps == 8 and stats.sorts == 4
    a, b = eager.get_state(), graph.get_state()
    for f in ("pos", "vel", "thq", "rho"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert np.array_equal(eager.get_rain_fx()[0], graph.get_rain_fx()[0])
    # a second replay reuses the captured graphs
    _eager(eager, dt, 4)
    graph.run_graph(4, dt, sort_every=2)
    a, b = eager.get_state(), graph.get_state()
    for f in ("pos", "vel", "thq", "rho"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
