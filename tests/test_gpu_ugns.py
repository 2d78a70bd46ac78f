"""UGNS on the GPU vs the FP64 oracle: sorted order, cell ranges and neighbour
lists must be bit-exact for identical FP32 positions (UGNSCudaImpl.cu:110-203)."""
import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes

pytestmark = pytest.mark.gpu


def _compare_ugns(wl, params=None):
    params = params or vtp.params_from_json(wl.solver, wl.config_json())
    ctx = vtp.Context(params)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, o.vol_frac, o.scene)
    ctx.initialize()
    _, cells = ctx.grid_info()
    S = wl.scenes()
    g = ctx.debug_neighbors(S)
    orc = pyoracle.Oracle.from_workload(params, wl)
    orc.initialize()
    r = orc.neighbors(cells * S)
    np.testing.assert_array_equal(g["sorted"], r["sorted"])
    np.testing.assert_array_equal(g["cell_start"], r["cell_start"])
    np.testing.assert_array_equal(g["cell_end"], r["cell_end"])
    np.testing.assert_array_equal(g["neighbor_num"], r["neighbor_num"])
    np.testing.assert_array_equal(g["neighbors"], r["neighbors"])
    st = ctx.stats()
    ctx.close()
    return g, st


def test_ugns_cfg0_positions_bit_exact():
    wl = scenes.cfg0_pbf()
    g, st = _compare_ugns(wl)
    assert st["dead_particles"] == 0
    assert g["neighbor_num"].max() <= 60


def test_ugns_random_points_bit_exact():
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.0, 0.3, size=(5000, 3)).astype(np.float32)
    wl = scenes.Workload("rand", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.3, 0.3, 0.3)),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    _compare_ugns(wl)


def test_ugns_dense_cap_truncation():
    # 400 particles inside one cell: lists are truncated at maxNeighborNum in
    # the reference traversal order.
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.1, 0.13, size=(400, 3)).astype(np.float32)
    wl = scenes.Workload("dense", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.4, 0.4, 0.4), max_nbr=20),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    g, _ = _compare_ugns(wl)
    assert (g["neighbor_num"] == 20).all()


def test_ugns_out_of_grid_and_duplicates():
    rng = np.random.default_rng(7)
    pos = rng.uniform(0.0, 0.2, size=(2000, 3)).astype(np.float32)
    pos[10] = [5.0, 0.1, 0.1]          # outside the grid
    pos[11] = [-0.5, 0.1, 0.1]         # outside (negative)
    pos[20] = pos[21]                  # coincident pair: distance 0 is not a neighbour
    wl = scenes.Workload("edge", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.2, 0.2, 0.2)),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    g, st = _compare_ugns(wl)
    assert st["dead_particles"] == 2
    assert g["neighbor_num"][10] == 0 and g["neighbor_num"][11] == 0
    assert 21 not in set(g["neighbors"][20, : g["neighbor_num"][20]])


def test_ugns_batched_scenes_do_not_mix():
    wl = scenes.cfg3_batched(scenes=4, lo=2000, hi=9000)
    g, _ = _compare_ugns(wl)
    sc = np.concatenate([np.full(o.pos.shape[0], o.scene) for o in wl.objects])
    lists = g["neighbors"]
    for p in range(0, lists.shape[0], 97):
        nb = lists[p, : g["neighbor_num"][p]]
        assert (sc[nb] == sc[p]).all()
