"""The benchmark workloads at their full sizes, checked through properties
that hold at any size (the oracle runs only the small cases):

* PBF dam break, 1,000,000 fluid + 321,440 wall particles: a graph replay and
  eager launches give the same ticks bit for bit, nothing leaves the grid or
  goes non-finite, the walls do not move;
* DFSPH dam break at 1M: every solve stops below its threshold within the
  iteration cap, no crash flag;
* IMM-CT three-phase at 1M (+591,065 walls, mass-weighted stress as the
  benchmark): every phase's total volume is
  conserved through ticks with negative-fraction retries (FP32 fractions
  summed in FP64), fractions stay a partition of one.
"""
import numpy as np
import pytest

import vtp
from vtp import scenes

pytestmark = pytest.mark.gpu


def _ctx(wl, **opts):
    params = vtp.params_from_json(wl.solver, wl.config_json())
    ctx = vtp.Context(params)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, o.vol_frac, o.scene)
    for k, v in opts.items():
        ctx.set_option(k, v)
    ctx.initialize()
    return ctx


def test_pbf_1m_graph_equals_eager_and_stays_on_the_grid():
    wl = scenes.pbf_1m(seed=0)
    a = _ctx(wl, graph=1)
    walls0 = a.read("pos")[wl.n_fluid:]
    b = _ctx(wl, graph=0)
    a.step(12)
    b.step(12)
    xa, xb = a.read("pos"), b.read("pos")
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(a.read("vel"), b.read("vel"))
    st = a.stats()
    assert st["dead_particles"] == 0 and st["crash_flags"] == 0
    assert np.isfinite(xa).all()
    np.testing.assert_array_equal(xa[wl.n_fluid:], walls0)
    a.close()
    b.close()


def test_dfsph_1m_solves_converge_within_the_cap():
    wl = scenes.dfsph_1m(seed=0)
    ctx = _ctx(wl, graph=1)
    req = wl.config["DFSPH"]["Required"]
    for _ in range(4):
        ctx.step(1)
        st = ctx.stats()
        assert st["crash_flags"] == 0 and st["dead_particles"] == 0
        assert 1 <= st["div_iterations"] <= 101 and 1 <= st["incomp_iterations"] <= 101
        assert st["div_residual"] < req["divFreeThreshold"]
        assert st["incomp_residual"] < req["incompThreshold"]
    ctx.close()


def test_immct_1m_conserves_every_phase():
    wl = scenes.cfg2_immct(seed=0)
    # as the benchmark: the reference's polymer-stress term (no neighbour mass)
    # blows up at this resolution within two ticks
    wl.config["IMMCT"]["Extensions"]["massWeighted"] = True
    ctx = _ctx(wl)
    nf = wl.n_fluid
    f0 = ctx.read("vol_frac")[:nf].astype(np.float64)
    s0 = f0.sum(0)
    rounds = []
    for _ in range(3):
        ctx.step(1)
        rounds.append(ctx.stats()["phase_iterations"])
    f = ctx.read("vol_frac")[:nf].astype(np.float64)
    np.testing.assert_allclose(f.sum(0), s0, rtol=2e-5)
    np.testing.assert_allclose(f.sum(1), 1.0, atol=1e-5)
    assert f.min() >= -1e-6
    assert ctx.stats()["crash_flags"] == 0 and max(rounds) >= 2
    ctx.close()
