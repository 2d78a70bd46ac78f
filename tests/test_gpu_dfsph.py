"""DFSPH step on the GPU vs the FP64 oracle (DFSPHCudaImpl.cu, DFSPHSolver.cpp:339-406).

Checks: iteration counts of both pressure solves equal the oracle's, exit
residuals agree (FP64 device reduction vs FP64 host sum, rel 1e-4 since the
per-particle deltas are FP32), and the fields after one step agree with the
oracle within 1e-5 (L2 norm-wise relative; max norm 1e-5 for x, v, m, V),
the acceleration (v* - v)/dt, the surface normal and the divergence solve's
kappa included.  The density solve's predictor delta = max(0, cr - 1 + dt
div v*) and its kappa are differences of numbers ~30x larger (cr ~ 1, delta
~ 0.03 in this compressed block), so the FP32 rounding of cr (~1e-7 relative,
measured 9.8e-8) reaches them amplified ~30-100x: 5e-5 (measured 1.1e-5).
"""
import numpy as np
import pytest

import pyoracle
import vtp
from conftest import rel_err
from vtp import scenes

pytestmark = pytest.mark.gpu

FIELD_TOL = {"pos": 1e-5, "vel": 1e-5, "mass": 1e-5, "volume": 1e-5, "rest_density": 1e-5,
             "compression_ratio": 1e-5, "df_alpha": 1e-5, "vel_adv": 1e-5, "surface_normal": 1e-5,
             "kappa_div": 1e-5, "acc": 1e-5, "delta_cr": 5e-5, "kappa_incomp": 5e-5}


def l2_rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb else float(np.linalg.norm(a))


def _pair(wl):
    params = vtp.params_from_json("DFSPH", wl.config_json())
    ctx = vtp.Context(params)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, None, o.scene)
    ctx.initialize()
    orc = pyoracle.Oracle.from_workload(params, wl)
    orc.initialize()
    return ctx, orc


def _compare(ctx, orc, fields=FIELD_TOL):
    ids = vtp._abi.FIELDS
    res = {}
    bad = {}
    for f, tol in fields.items():
        g = ctx.read(f)
        r = orc.read(ids[f][0])
        e = l2_rel(g, r)
        res[f] = e
        if not e <= tol:
            bad[f] = e
    assert not bad, f"over tolerance: {bad} (all {res})"
    return res


def test_dfsph_init_matches():
    wl = scenes.cfg1_dfsph(n_side=(16, 16, 16))
    ctx, orc = _pair(wl)
    _compare(ctx, orc, {"pos": 1e-7, "vel": 1e-7, "mass": 1e-6, "volume": 1e-6, "rest_density": 1e-6})


@pytest.mark.parametrize("tension", [0.0, 0.05])
def test_dfsph_one_step(tension):
    wl = scenes.cfg1_dfsph(n_side=(24, 24, 24))
    wl.config["DFSPH"]["Required"]["surfaceTensionCoefficient"] = tension
    ctx, orc = _pair(wl)
    ctx.step(1)
    orc.step(1)
    gs, rs = ctx.stats(), orc.stats()
    assert gs["div_iterations"] == rs["div_iterations"]
    assert gs["incomp_iterations"] == rs["incomp_iterations"]
    assert abs(gs["div_residual"] - rs["div_residual"]) <= 1e-4 * abs(rs["div_residual"]) + 1e-12
    assert abs(gs["incomp_residual"] - rs["incomp_residual"]) <= 1e-4 * abs(rs["incomp_residual"]) + 1e-12
    _compare(ctx, orc)
    for f in ("pos", "vel", "mass", "volume"):
        assert rel_err(ctx.read(f), orc.read(vtp._abi.FIELDS[f][0])) <= 1e-5


def test_dfsph_five_steps_iteration_counts():
    wl = scenes.cfg1_dfsph(n_side=(20, 20, 20))
    ctx, orc = _pair(wl)
    for _ in range(5):
        ctx.step(1)
        orc.step(1)
        gs, rs = ctx.stats(), orc.stats()
        assert (gs["div_iterations"], gs["incomp_iterations"]) == (rs["div_iterations"], rs["incomp_iterations"])
    _compare(ctx, orc, {"pos": 1e-5, "vel": 1e-4})


def test_dfsph_iteration_cap_sets_crash():
    wl = scenes.cfg1_dfsph(n_side=(12, 12, 12))
    wl.config["DFSPH"]["Required"]["incompThreshold"] = 0.0   # never converges
    wl.config["DFSPH"]["Extensions"] = {"maxSolverIterations": 7}
    ctx, orc = _pair(wl)
    with pytest.raises(vtp.VtpError, match="VTP_ECRASH"):
        ctx.step(1)
    st = ctx.stats()
    assert st["incomp_iterations"] == 8          # cap + 1, as iter > cap breaks
    assert st["crash_flags"] & vtp._abi.CRASH_INCOMP_ITER_CAP
    orc.step(1)
    assert orc.stats()["incomp_iterations"] == 8
    with pytest.raises(vtp.VtpError, match="crashed"):
        ctx.step(1)


def test_dfsph_cfl_step_closed_form():
    """CFL extension: dt = min(dtMax, cfl * 2r / max|v_fluid|), measured at the
    start of the tick.  An uncompressed cluster translating at 8 m/s with
    gravity off feels no pressure or viscosity, so it moves by exactly u*dt."""
    import json
    r = 0.01
    cfg = scenes.dfsph_config(radius=r, lb=(0, 0, 0), size=(1, 1, 1), dt=0.002, cfl=True)
    cfg["DFSPH"]["Optional"] = {"enable": True, "gravity": [0.0, 0.0, 0.0]}
    s = 1.1 * 2 * r
    g = np.stack(np.meshgrid(*(np.arange(6) * s,) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32) + 0.3
    u = np.array([8.0, 0.0, 0.0], np.float32)
    p = vtp.params_from_json("DFSPH", json.dumps(cfg))
    ctx = vtp.Context(p)
    ctx.add_particles(g, vtp.MAT_FLUID, np.tile(u, (g.shape[0], 1)))
    ctx.initialize()
    ctx.step(1)
    dt = np.float32(np.float32(0.4) * np.float32(2 * r) / np.float32(8.0))
    st = ctx.stats()
    assert st["dt"] == pytest.approx(float(dt), rel=1e-6)
    assert st["sim_time"] == pytest.approx(float(dt), rel=1e-6)
    np.testing.assert_allclose(ctx.read("pos"), g + u * dt, atol=2e-7)
    ctx.close()


def test_dfsph_two_fluid_objects_weight_the_residual_per_object():
    """The solve residual is the mean over the fluid objects of each object's
    mean delta (CommonFunc.cpp:10-21), so particles of a smaller object weigh
    more.  Two fluid blocks of 8,000 and 1,000 particles (the small one
    compressed by a 0.8x spacing, so the two objects' mean deltas differ)
    plus the walls, 3 ticks: iteration counts equal the oracle's and both
    exit residuals agree within 1e-4 every tick."""
    s = 0.01
    big = scenes.jitter(scenes.particle_cube(s / 2, (0.02, 0.02, 0.02), (20 * s, 20 * s, 20 * s)), s, 1)
    small = scenes.jitter(scenes.particle_cube(0.4 * s, (1.5, 0.02, 0.02), (8 * s, 8 * s, 8 * s)), 0.8 * s, 2)
    walls = scenes.particle_box(s / 2, (0.0, 0.0, 0.0), (2.0, 1.0, 1.0), 2)
    assert big.shape[0] == 8000 and small.shape[0] == 1000
    wl = scenes.Workload("two_blocks", "DFSPH", scenes.dfsph_config(),
                         [scenes.Obj(vtp.MAT_FLUID, big), scenes.Obj(vtp.MAT_FLUID, small),
                          scenes.Obj(vtp.MAT_BOUNDARY, walls)])
    ctx, orc = _pair(wl)
    for t in range(3):
        ctx.step(1)
        orc.step(1)
        gs, rs = ctx.stats(), orc.stats()
        print(f"tick {t + 1}: iterations {gs['div_iterations']}, {gs['incomp_iterations']}; residuals "
              f"{gs['div_residual']:.6e} / {rs['div_residual']:.6e}, {gs['incomp_residual']:.6e} / "
              f"{rs['incomp_residual']:.6e}", flush=True)
        assert (gs["div_iterations"], gs["incomp_iterations"]) == (rs["div_iterations"], rs["incomp_iterations"])
        assert abs(gs["div_residual"] - rs["div_residual"]) <= 1e-4 * abs(rs["div_residual"]) + 1e-12
        assert abs(gs["incomp_residual"] - rs["incomp_residual"]) <= 1e-4 * abs(rs["incomp_residual"]) + 1e-12
    ctx.close()
