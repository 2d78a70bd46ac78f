"""Batched scenes: several independent scenes in one context (segmented
keys scene*cells + cell, one launch per stage) give, scene by scene, exactly
the state a context holding that scene alone reaches."""
import numpy as np
import pytest

import vtp
from vtp import scenes

pytestmark = pytest.mark.gpu


def test_batched_pbf_equals_one_context_per_scene():
    wl = scenes.cfg3_batched(seed=1, scenes=5, lo=2000, hi=9000)
    params = vtp.params_from_json("PBF", wl.config_json())
    big = vtp.Context(params)
    for o in wl.objects:
        big.add_particles(o.pos, o.mat, o.vel, None, o.scene)
    big.set_option("graph", 1)
    big.initialize()
    big.step(6)
    xb, vb = big.read("pos"), big.read("vel")
    start = 0
    per = {}
    for o in wl.objects:
        per.setdefault(o.scene, []).append(o)
    for sc, objs in sorted(per.items()):
        one = vtp.Context(params)
        for o in objs:
            one.add_particles(o.pos, o.mat, o.vel, None, 0)
        one.initialize()
        one.step(6)
        n = sum(o.pos.shape[0] for o in objs)
        np.testing.assert_array_equal(one.read("pos"), xb[start:start + n])
        np.testing.assert_array_equal(one.read("vel"), vb[start:start + n])
        start += n
        one.close()
    assert start == wl.n
    big.close()


def _l2_rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("graph", [0, 1])
def test_batched_pbf_tick_matches_oracle(graph):
    """The batched tick (one context, segmented keys scene*cells + cell, one
    launch per stage, optionally replayed from a hipGraph) against the FP64
    oracle run with the same scene ids: one tick within 1e-5 (x, v, density;
    C and lambda 5e-5, the cancellation bound of the single-scene tests),
    then five more (positions 1e-5, velocities 1e-4, L2-relative), for the
    whole batch and for each scene on its own."""
    import pyoracle
    wl = scenes.cfg3_batched(seed=2, scenes=6, lo=1500, hi=6000)
    params = vtp.params_from_json("PBF", wl.config_json())
    ctx = vtp.Context(params)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, None, o.scene)
    ctx.set_option("graph", graph)
    ctx.initialize()
    orc = pyoracle.Oracle.from_workload(params, wl)
    orc.initialize()
    ids = vtp._abi.FIELDS
    print(f"batched PBF tick vs oracle: {wl.scenes()} scenes, {wl.n} particles, graph={graph}", flush=True)
    ctx.step(1)
    orc.step(1)
    tol = {"pos": 1e-5, "vel": 1e-5, "density": 1e-5, "error": 5e-5, "lambda": 5e-5}
    errs = {f: _l2_rel(ctx.read(f), orc.read(ids[f][0])) for f in tol}
    assert all(errs[f] <= tol[f] for f in tol), errs
    ctx.step(5)
    orc.step(5)
    sc = ctx.read("scene")
    x, xr = ctx.read("pos"), orc.read(0)
    v, vr = ctx.read("vel"), orc.read(1)
    assert _l2_rel(x, xr) <= 1e-5 and _l2_rel(v, vr) <= 1e-4
    for s in range(wl.scenes()):
        m = sc == s
        assert _l2_rel(x[m], xr[m]) <= 1e-5, s
        assert _l2_rel(v[m], vr[m]) <= 1e-4, s
    ctx.close()
