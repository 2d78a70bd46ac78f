"""UGNS on the GPU vs the FP64 oracle: sorted order, cell ranges and neighbour
lists must be bit-exact for identical FP32 positions (UGNSCudaImpl.cu:110-203)."""
import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes

pytestmark = pytest.mark.gpu


def _compare_ugns(wl, params=None, walk=None, radix=None):
    params = params or vtp.params_from_json(wl.solver, wl.config_json())
    ctx = vtp.Context(params)
    if walk is not None:
        ctx.set_option("neighbor_walk", walk)
    if radix is not None:
        ctx.set_option("radix_sort", radix)
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


@pytest.mark.parametrize("walk,radix", [(1, 0), (2, 0), (1, 1)])
def test_ugns_random_points_bit_exact(walk, radix):
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.0, 0.3, size=(5000, 3)).astype(np.float32)
    wl = scenes.Workload("rand", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.3, 0.3, 0.3)),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    _compare_ugns(wl, walk=walk, radix=radix)


@pytest.mark.parametrize("walk", [1, 2])
@pytest.mark.parametrize("max_nbr", [2, 5, 9, 17, 20, 33, 34, 41, 49])
def test_ugns_dense_cap_truncation(walk, max_nbr):
    # 400 particles inside one cell: lists are truncated at maxNeighborNum in
    # the reference traversal order.  Every list then holds exactly cap =
    # maxNeighborNum - 1 entries, so the caps cover a list that ends inside,
    # or exactly at the end of, an 8-entry chunk and one that runs past the
    # build's 32-entry LDS stage into entries stored directly.
    rng = np.random.default_rng(5)
    pos = rng.uniform(0.1, 0.13, size=(400, 3)).astype(np.float32)
    wl = scenes.Workload("dense", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.4, 0.4, 0.4), max_nbr=max_nbr),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    g, _ = _compare_ugns(wl, walk=walk)
    assert (g["neighbor_num"] == max_nbr).all()


@pytest.mark.parametrize("radix", [0, 1])
def test_ugns_out_of_grid_and_duplicates(radix):
    rng = np.random.default_rng(7)
    pos = rng.uniform(0.0, 0.2, size=(2000, 3)).astype(np.float32)
    pos[10] = [5.0, 0.1, 0.1]          # outside the grid
    pos[11] = [-0.5, 0.1, 0.1]         # outside (negative)
    pos[20] = pos[21]                  # coincident pair: distance 0 is not a neighbour
    wl = scenes.Workload("edge", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.2, 0.2, 0.2)),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    g, st = _compare_ugns(wl, radix=radix)
    assert st["dead_particles"] == 2
    assert g["neighbor_num"][10] == 0 and g["neighbor_num"][11] == 0
    assert 21 not in set(g["neighbors"][20, : g["neighbor_num"][20]])


@pytest.mark.parametrize("radix", [0, 1])
def test_ugns_batched_scenes_do_not_mix(radix):
    wl = scenes.cfg3_batched(scenes=4, lo=2000, hi=9000)
    g, _ = _compare_ugns(wl, radix=radix)
    sc = np.concatenate([np.full(o.pos.shape[0], o.scene) for o in wl.objects])
    lists = g["neighbors"]
    for p in range(0, lists.shape[0], 97):
        nb = lists[p, : g["neighbor_num"][p]]
        assert (sc[nb] == sc[p]).all()


@pytest.mark.parametrize("walk", [1, 2])
def test_ugns_cap_cut_inside_the_x_plus_one_cells(walk):
    # Cell (1,2,2) holds 10 particles, cell (2,2,2) 60, all within h of each
    # other.  A particle of the first cell finds 9 neighbours in the x-1 and x
    # cells and 60 in the x+1 cell: its list (cap 19) keeps the first 10 of
    # the x+1 cell in the reference order.  A particle of the second cell
    # finds 69 in its x-1 and x cells: the reference traversal decides.
    rng = np.random.default_rng(11)
    a = np.column_stack([rng.uniform(0.070, 0.079, 10), rng.uniform(0.10, 0.11, (10, 2))])
    b = np.column_stack([rng.uniform(0.081, 0.090, 60), rng.uniform(0.10, 0.11, (60, 2))])
    pos = np.concatenate([a, b]).astype(np.float32)
    wl = scenes.Workload("xcut", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.4, 0.4, 0.4), max_nbr=20),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    g, _ = _compare_ugns(wl, walk=walk)
    assert (g["neighbor_num"] == 20).all()
    first = g["neighbors"][0, 1:20]
    assert (first < 10).sum() == 9 and (first >= 10).sum() == 10


@pytest.mark.parametrize("walk", [1, 2])
def test_ugns_density_gradient_bit_exact(walk):
    # lattice-free points whose density rises 27-fold towards x = 0, so lists
    # range from a few entries to far past the cap
    rng = np.random.default_rng(13)
    u = rng.uniform(0.0, 1.0, size=(3000, 3))
    pos = np.column_stack([0.3 * u[:, 0] ** 3, 0.3 * u[:, 1], 0.3 * u[:, 2]]).astype(np.float32)
    wl = scenes.Workload("grad", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.3, 0.3, 0.3), max_nbr=40),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    g, _ = _compare_ugns(wl, walk=walk)
    num = g["neighbor_num"]
    assert (num == 40).mean() > 0.2 and (num < 40).mean() > 0.2


@pytest.mark.parametrize("walk", [1, 2])
def test_ugns_cap_reached_in_the_x_minus_one_and_x_cells(walk):
    # 20 particles in cell (1,2,2) and 10 in cell (2,2,2), all within h.  With
    # cap 19, a particle of the first cell fills its list from its own cell
    # exactly (19 = cap); one of the second cell meets 29 in its x-1 and x
    # cells and is rebuilt by the reference traversal.
    rng = np.random.default_rng(17)
    a = np.column_stack([rng.uniform(0.075, 0.0799, 20), rng.uniform(0.10, 0.105, (20, 2))])
    b = np.column_stack([rng.uniform(0.0801, 0.085, 10), rng.uniform(0.10, 0.105, (10, 2))])
    pos = np.concatenate([a, b]).astype(np.float32)
    wl = scenes.Workload("xfront", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.4, 0.4, 0.4), max_nbr=20),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    g, _ = _compare_ugns(wl, walk=walk)
    assert (g["neighbor_num"] == 20).all()
    assert (g["neighbors"][0, 1:20] < 20).all()
    assert (g["neighbors"][25, 1:20] < 20).all()


def _shell_points(h, lb, seed):
    """Anchors on and next to cell faces, edges and corners, each with partners
    at distances just below, at and just above h in axis, face-diagonal,
    body-diagonal and random directions."""
    rng = np.random.default_rng(seed)
    hf = float(np.float32(h))
    eps = float(np.spacing(np.float32(hf)))
    offs = [0.0, eps, 0.5 * hf, hf - eps]
    dirs = []
    for v in [(1, 0, 0), (0, 1, 0), (0, 0, 1)]:
        dirs += [v, tuple(-c for c in v)]
    for sx in (1, -1):
        for sy in (1, -1):
            dirs += [(sx, sy, 0), (sx, 0, sy), (0, sx, sy)]
            for sz in (1, -1):
                dirs.append((sx, sy, sz))
    dirs = [np.array(d, float) / np.linalg.norm(d) for d in dirs]
    dirs += list(rng.normal(size=(10, 3)))
    dirs = [d / np.linalg.norm(d) for d in dirs]
    pts = []
    for i in (2, 5, 8):
        for j in (2, 5, 8):
            for k in (2, 5, 8):
                o = rng.integers(0, len(offs), 3)
                a = np.array([lb + (i + 0.0) * hf + offs[o[0]], lb + j * hf + offs[o[1]], lb + k * hf + offs[o[2]]])
                a = a.astype(np.float32).astype(float)
                pts.append(a)
                for d in (hf * (1 - 2.0 ** -20), hf, hf * (1 + 2.0 ** -20)):
                    for u in dirs:
                        pts.append(a + d * u)
    return np.array(pts, dtype=np.float32)


@pytest.mark.parametrize("walk", [1, 2])
def test_ugns_pairs_at_the_cutoff_across_cell_faces_and_corners(walk):
    # Pairs at distance h (and one float either side) whose partner sits in a
    # face, edge or corner cell.  Each is decided by the FP32 test with every
    # product rounded (a fused dx*dx + dy*dy moves pairs across the cut-off),
    # and no pair in an edge or corner cell may be lost.
    pos = _shell_points(0.04, 0.0, seed=19)
    wl = scenes.Workload("shell", "PBF", scenes.pbf_config(radius=0.01, lb=(0, 0, 0), size=(0.5, 0.5, 0.5),
                                                           max_nbr=256),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    g, st = _compare_ugns(wl, walk=walk)
    assert st["dead_particles"] == 0
    assert g["neighbor_num"].max() < 256


def _per_tick_workload(solver):
    """A jittered 12^3 block in its box, four isolated particles just outside
    the box's +z walls (still on the grid) and two particles off the grid
    from the start.  In the DFSPH run the four move out at 5, 8, 12 and
    16 m/s and have left the grid by the searches of ticks 4, 3, 3 and 2; the
    PBF run zeroes the velocities before every tick, so there they stay."""
    if solver == "PBF":
        cfg = scenes.pbf_config(lb=(-0.1, -0.1, -0.1), size=(0.6, 0.6, 0.6))
        cfg["PBF"]["Optional"] = {"enable": True, "gravity": [0.0, 0.0, 0.0]}
        spacing, zfly = 0.02, 0.485
    else:
        cfg = scenes.dfsph_config(lb=(-0.05, -0.05, -0.05), size=(0.3, 0.3, 0.3))
        spacing, zfly = 0.01, 0.2425
    objs = scenes.dam_break("tick", solver, cfg, (12, 12, 12), spacing, (spacing * 20,) * 3,
                            (spacing * 3,) * 3, seed=4)
    fly = np.array([[x, 0.5, zfly / (spacing * 20)] for x in (0.15, 0.35, 0.55, 0.75)], np.float32) * spacing * 20
    fly_v = np.zeros((4, 3), np.float32)
    if solver == "DFSPH":
        fly_v[:, 2] = [5.0, 8.0, 12.0, 16.0]
    gone = np.array([[50.0, 0.0, 0.0], [-50.0, 0.1, 0.1]], np.float32)
    objs.append(scenes.Obj(vtp.MAT_FLUID, fly, fly_v))
    objs.append(scenes.Obj(vtp.MAT_FLUID, gone))
    return scenes.Workload("per_tick", solver, cfg, objs)


@pytest.mark.parametrize("solver", ["PBF", "DFSPH"])
@pytest.mark.parametrize("walk,radix", [(1, 0), (2, 0), (1, 1), (2, 1)])
def test_ugns_per_tick_search_bit_exact(solver, walk, radix):
    """The search every tick runs (UGNSCudaImpl.cu:205-227) starts from the
    previous tick's slot order and restores the (cellId, attach id) order in
    the cells (k_cell_fixup); its lists are built after the state is
    reordered.  Before each of four ticks the block is shifted by
    +-(0.3, 0.2, 0.1) h, so a third of it changes cells every tick; after the
    tick the device's sorted ids, cell tables and neighbour lists (reference
    layout) equal the oracle's search of the same FP32 positions: the
    positions the tick searched.

    DFSPH searches the positions it starts the tick with.  PBF searches its
    predicted positions; with gravity off and the velocities zeroed before a
    tick, x + dt (v + dt g) is x bit for bit, so the written positions are
    the searched ones.  Off-grid particles sit after the live ones; the
    reference leaves their order to stale data (A10), so they are compared as
    a set.
    """
    wl = _per_tick_workload(solver)
    params = vtp.params_from_json(solver, wl.config_json())
    ctx = vtp.Context(params)
    ctx.set_option("neighbor_walk", walk)
    ctx.set_option("radix_sort", radix)
    if solver in ("PBF", "DFSPH"):
        ctx.set_option("static_lists", 1)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, o.vol_frac, o.scene)
    ctx.initialize()
    _, cells = ctx.grid_info()
    h = 4 * params.particle_radius
    block = np.arange(wl.objects[0].pos.shape[0])
    orc = pyoracle.Oracle.from_workload(params, wl)
    deads = []
    moved = []
    prev = None
    for t, sign in enumerate((1, 1, -1, -1)):
        pos = ctx.read("pos")
        pos[block] += (sign * np.array([0.3, 0.2, 0.1]) * h).astype(np.float32)
        ctx.write("pos", pos)
        if solver == "PBF":
            ctx.write("vel", np.zeros((wl.n, 3), np.float32))
        print(f"per-tick search {solver} walk={walk} radix={radix} tick {t + 1}", flush=True)
        ctx.step(1)
        g = ctx.debug_neighbors(1)
        orc.write(0, pos.astype(np.float64))
        orc.search()
        r = orc.neighbors(cells)
        dead = int((r["neighbor_num"] == 0).sum())   # a live particle lists itself
        assert ctx.stats()["dead_particles"] == dead
        deads.append(dead)
        live = wl.n - dead
        np.testing.assert_array_equal(g["sorted"][:live], r["sorted"][:live])
        assert set(g["sorted"][live:].tolist()) == set(r["sorted"][live:].tolist())
        np.testing.assert_array_equal(g["cell_start"], r["cell_start"])
        np.testing.assert_array_equal(g["cell_end"], r["cell_end"])
        np.testing.assert_array_equal(g["neighbor_num"], r["neighbor_num"])
        np.testing.assert_array_equal(g["neighbors"], r["neighbors"])
        if prev is not None:
            moved.append(int((g["sorted"] != prev).sum()))
        prev = g["sorted"].copy()
    # slots changed between ticks; particles left the grid during the DFSPH run
    assert min(moved) > wl.n // 10, moved
    assert deads == ([2, 2, 2, 2] if solver == "PBF" else [2, 3, 5, 6]), deads
    ctx.close()


def _layer_spans(cs, ce, grid, keys):
    """Largest slot span of one cell layer's three rows (x-1..x+1) over the
    particles with the given cell keys: the offsets a 16-bit list entry holds."""
    gx, gy, gz = grid
    worst = 0
    for k in np.unique(keys):
        cz, rem = divmod(int(k), gx * gy)
        cy, cx = divmod(rem, gx)
        for z in (cz - 1, cz, cz + 1):
            if not 0 <= z < gz:
                continue
            lo, hi = None, None
            for y in (cy - 1, cy, cy + 1):
                if not 0 <= y < gy:
                    continue
                for x in (cx - 1, cx, cx + 1):
                    if not 0 <= x < gx:
                        continue
                    c = (z * gy + y) * gx + x
                    if cs[c] != 0xFFFFFFFF:
                        lo = cs[c] if lo is None else lo
                        hi = ce[c]
            if lo is not None:
                worst = max(worst, int(hi) - int(lo))
    return worst


@pytest.mark.parametrize("walk", [1, 2])
def test_ugns_wide_lists_bit_exact(walk):
    """A sheet a few cell rows high and 4,000 cells long: the rows of one
    cell layer lie ~24k slots apart, more than a 14-bit list offset holds, so
    those 64-slot groups store 32-bit entries (the fallback build).  The
    lists (reference layout) equal the oracle's, and three PBF ticks through
    the wide lists match the oracle's."""
    s = 0.02
    nx, ny, nz = 8000, 4, 2
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pos = np.stack([0.03 + s * ix, 0.03 + s * iy, 0.03 + s * iz], -1).reshape(-1, 3)
    pos = scenes.jitter(pos.astype(np.float32), s, seed=12)
    cfg = scenes.pbf_config(lb=(0, 0, 0), size=(nx * s + 0.2, 0.2, 0.16))
    cfg["PBF"]["Optional"] = {"enable": True, "gravity": [0.0, 0.0, 0.0]}
    wl = scenes.Workload("sheet", "PBF", cfg, [scenes.Obj(vtp.MAT_FLUID, pos)])
    params = vtp.params_from_json("PBF", wl.config_json())
    ctx = vtp.Context(params)
    ctx.set_option("neighbor_walk", walk)
    ctx.set_option("static_lists", 1)
    ctx.add_particles(pos, vtp.MAT_FLUID)
    ctx.initialize()
    grid, cells = ctx.grid_info()
    orc = pyoracle.Oracle.from_workload(params, wl)
    orc.initialize()
    r = orc.neighbors(cells)
    # the span that decides the format, from the oracle's cell tables
    gx, gy, gz = grid
    cell_of = lambda p: ((np.floor(p[:, 2] / (4 * params.particle_radius)).astype(np.int64) * gy +
                          np.floor(p[:, 1] / (4 * params.particle_radius)).astype(np.int64)) * gx +
                         np.floor(p[:, 0] / (4 * params.particle_radius)).astype(np.int64))
    keys = cell_of(pos.astype(np.float64))
    assert _layer_spans(r["cell_start"], r["cell_end"], grid, keys[::7]) > (1 << 14)
    print(f"wide lists walk={walk}: {pos.shape[0]} particles", flush=True)
    g = ctx.debug_neighbors(1)
    np.testing.assert_array_equal(g["neighbor_num"], r["neighbor_num"])
    np.testing.assert_array_equal(g["neighbors"], r["neighbors"])
    for _ in range(3):
        ctx.step(1)
        orc.step(1)
    g, r = ctx.debug_neighbors(1), orc.neighbors(cells)
    np.testing.assert_array_equal(g["neighbors"], r["neighbors"])
    x, xr = ctx.read("pos"), orc.read(0)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) <= 1e-5
    ctx.close()
