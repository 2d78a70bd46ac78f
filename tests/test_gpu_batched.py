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
