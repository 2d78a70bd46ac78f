"""The synthetic benchmark inputs against the committed fixture
(tests/golden/workload_heads.json, written by tests/golden/make_fixtures.py):
particle counts, the first 16 positions and velocities bit for bit, and the
FP64 sum of all positions."""
import json
from pathlib import Path

import numpy as np
import pytest

from vtp import scenes

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "workload_heads.json").read_text())
MAKE = {
    "cfg0_pbf": lambda: scenes.cfg0_pbf(seed=0),
    "cfg1_dfsph": lambda: scenes.cfg1_dfsph(seed=0),
    "cfg2_immct": lambda: scenes.cfg2_immct(seed=0),
    "cfg3_batched": lambda: scenes.cfg3_batched(seed=0),
    "cfg4_adjoint": lambda: scenes.cfg4_adjoint(seed=0),
    "pbf_1m": lambda: scenes.pbf_1m(seed=0),
    "dfsph_1m": lambda: scenes.dfsph_1m(seed=0),
}


@pytest.mark.parametrize("name", sorted(MAKE))
def test_workload_matches_fixture(name):
    g = GOLDEN[name]
    wl = MAKE[name]()
    pos = np.concatenate([o.pos for o in wl.objects]).astype(np.float32)
    vel = np.concatenate([o.vel if o.vel is not None else np.zeros_like(o.pos) for o in wl.objects])
    assert (pos.shape[0], wl.n_fluid, len(wl.objects)) == (g["n"], g["n_fluid"], g["objects"])
    np.testing.assert_array_equal(pos[:16], np.asarray(g["pos"], np.float32))
    np.testing.assert_array_equal(vel[:16].astype(np.float32), np.asarray(g["vel"], np.float32))
    assert pos.astype(np.float64).sum() == g["pos_sum"]
