"""Checkpoint / resume through the C ABI (vtp_state_size, vtp_save_state,
vtp_load_state).  The reference has no resume path (ExportUtil.cpp:28-88
writes positions and colour only), so the bar is self-consistency, bit for
bit: a context that loads a state image and ticks K times ends in exactly the
state the saved context reaches after the same K ticks, for every carried
field, iteration count and the simulated time.
"""
import ctypes

import numpy as np
import pytest

import vtp
from vtp import scenes

pytestmark = pytest.mark.gpu

CARRIED = {
    "PBF": ["pos", "vel", "mass", "volume"],
    "DFSPH": ["pos", "vel", "mass", "volume"],
    "IMM": ["pos", "vel", "mass", "volume", "vol_frac", "vel_phase", "rest_density", "color"],
    "IMMCT": ["pos", "vel", "mass", "volume", "vol_frac", "vel_phase", "rest_density", "color", "Q"],
}


def _workload(solver):
    if solver == "PBF":
        return scenes.cfg0_pbf(seed=3, n_side=(16, 16, 16))
    if solver == "DFSPH":
        return scenes.cfg1_dfsph(seed=3, n_side=(16, 16, 16), cfl=True)
    return scenes.cfg2_immct(seed=3, n_side=(12, 12, 12), box=(0.3, 0.3, 0.3), solver=solver)


def _context(wl, solver):
    ctx = vtp.Context(vtp.params_from_json(solver, wl.config_json()))
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, o.vol_frac, o.scene)
    ctx.initialize()
    return ctx


def _snapshot(ctx, fields):
    s = {f: ctx.read(f).copy() for f in fields}
    st = ctx.stats()
    s["_stats"] = {k: st[k] for k in ("steps", "sim_time", "div_iterations", "incomp_iterations",
                                      "phase_iterations", "total_div_iterations", "total_incomp_iterations",
                                      "crash_flags")}
    return s


def _assert_same(a, b):
    for k in a:
        if k == "_stats":
            assert a[k] == b[k], (a[k], b[k])
        else:
            assert a[k].dtype == b[k].dtype and a[k].shape == b[k].shape, k
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), f"{k} differs after resume"


@pytest.mark.parametrize("solver", ["PBF", "DFSPH", "IMM", "IMMCT"])
def test_resume_is_bit_exact(solver):
    print(f"[state] resume {solver}", flush=True)
    wl = _workload(solver)
    fields = CARRIED[solver]
    a = _context(wl, solver)
    a.step(6)
    image = a.save_state()
    h, arrays = vtp.parse_state(image)
    assert h.n == a.n and h.steps == 6 and h.solver == vtp._abi.SOLVER_IDS[solver]
    assert np.array_equal(np.sort(arrays["orig"]), np.arange(a.n, dtype=np.uint32))
    # the image's slot-order positions are the attach-order field, permuted
    pos = a.read("pos")
    assert np.array_equal(arrays["x"][:, :3], pos[arrays["orig"]])
    a.step(5)
    want = _snapshot(a, fields)

    b = _context(wl, solver)
    b.step(2)                 # unrelated progress that the load must replace
    b.load_state(image)
    assert b.stats()["steps"] == 6
    b.step(5)
    _assert_same(want, _snapshot(b, fields))
    a.close()
    b.close()


def test_reload_same_context_and_file_roundtrip(tmp_path):
    print("[state] reload PBF", flush=True)
    wl = _workload("PBF")
    a = _context(wl, "PBF")
    a.set_option("graph", 1)
    a.step(3)
    path = tmp_path / "pbf.vtpstate"
    image = a.save_state(path)
    assert path.read_bytes() == image
    a.step(4)
    want = _snapshot(a, CARRIED["PBF"])
    a.load_state(path)        # rewind the same context (graph replay keeps its buffers)
    with pytest.raises(vtp.VtpError):
        a.debug_neighbors()   # lists belong to the replaced state until the next tick
    a.step(4)
    _assert_same(want, _snapshot(a, CARRIED["PBF"]))
    a.close()


def test_load_rejects_other_layouts():
    print("[state] rejects", flush=True)
    wl = _workload("PBF")
    a = _context(wl, "PBF")
    a.step(1)
    image = a.save_state()
    before = a.read("pos").copy()
    small = _context(scenes.cfg0_pbf(seed=3, n_side=(8, 8, 8)), "PBF")
    with pytest.raises(vtp.VtpError, match="objects differ"):
        small.load_state(image)
    with pytest.raises(vtp.VtpError, match="size"):
        a.load_state(image[:-4])
    bad = bytearray(image)
    bad[:8] = b"NOTSTATE"
    with pytest.raises(vtp.VtpError, match="not a state image"):
        a.load_state(bytes(bad))
    # a slot map that is not a permutation never reaches the device
    off = ctypes.sizeof(vtp.VtpStateHeader)
    dup = bytearray(image)
    dup[off + 4:off + 8] = dup[off:off + 4]
    with pytest.raises(vtp.VtpError, match="permutation"):
        a.load_state(bytes(dup))
    assert np.array_equal(a.read("pos"), before)   # failed loads change nothing
    dfsph = _context(_workload("DFSPH"), "DFSPH")
    with pytest.raises(vtp.VtpError, match="solver type"):
        dfsph.load_state(image)
    for c in (a, small, dfsph):
        c.close()


def test_resume_keeps_the_adaptive_list_walk():
    # Most lists of this scene reach the cap, so after the first tick every
    # build takes the reference walk, whose list order (and FP32 rounding)
    # differs from the row walk.  The image carries the cap count, so the
    # resumed context takes the same walks as the saved one.
    print("[state] adaptive walk", flush=True)
    rng = np.random.default_rng(13)
    u = rng.uniform(0.0, 1.0, size=(3000, 3))
    pos = np.column_stack([0.3 * u[:, 0] ** 3, 0.3 * u[:, 1], 0.3 * u[:, 2]]).astype(np.float32)
    wl = scenes.Workload("grad", "PBF", scenes.pbf_config(lb=(0, 0, 0), size=(0.3, 0.3, 0.3), max_nbr=40),
                         [scenes.Obj(vtp.MAT_FLUID, pos)])
    a = _context(wl, "PBF")
    a.step(1)
    image = a.save_state()
    h, _ = vtp.parse_state(image)
    assert h.nbr_at_cap > h.n // 8   # about half the lists: the next build takes the reference walk
    a.step(3)
    want = _snapshot(a, CARRIED["PBF"])
    b = _context(wl, "PBF")
    b.load_state(image)
    b.step(3)
    _assert_same(want, _snapshot(b, CARRIED["PBF"]))
    a.close()
    b.close()
