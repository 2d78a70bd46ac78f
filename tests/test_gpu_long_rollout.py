"""Long rollouts: statistics that must hold over hundreds of ticks.

* DFSPH dam break, 64^3 fluid at r = 0.005 with the CFL time step (cflNumber
  0.4, dtMax 0.002), 1000 ticks on the GPU.  Every divergence and density
  solve must stop because its residual fell below the threshold
  (DFSPHCudaImpl.cu:449-503, 556-611: the loop exits when the residual is
  under the threshold or the iteration count passes maxSolverIterations,
  so at most 101 iterations; the cap must never be the exit here, which the
  crash word records), no particle may leave the grid, and the fluid may not
  compress: over the fluid particles whose compression ratio is at least
  0.9 (the bulk, not the free surface), the mean ratio stays within 1 +/- 0.01
  and its 99.9th percentile under 1.05 at every 100th tick.
* PBF dam break, 32^3 fluid (config-0 physics), 500 ticks on the GPU and in
  the FP64 oracle.  A dam break is chaotic over 500 ticks, so the particle
  states part ways; the statistics may not.  Sampled every 10 ticks and
  averaged over 50-tick windows, over the fluid still on the grid: its
  count, its mean density, the mean and the 99th percentile of the density
  constraint C and its SPH volume (sum of m / rho).  With the reference's
  constraint gradient (no neighbour mass) the correction is weak at this
  particle mass and the column sinks through the two-layer floor and leaves
  the grid by tick ~250, in the oracle as on the GPU, so the run is also
  made with the mass-weighted gradient (extension massWeighted), where the
  column settles in the box (about a dozen particles slip out through the
  wall layers over the 500 ticks, in the oracle as on the GPU).  Their bounds come from the spread of the same
  window means among the oracle and three oracle runs whose fluid started
  1e-7 m away (a perturbation at the FP32 rounding of the positions, which
  is what separates the GPU run from the oracle): each GPU window mean must
  lie within 3x the largest such spread of the oracle's, plus a floor
  (5e-5 relative for density and volume, 5e-5 for the mean of C, 2e-3 for
  its percentile, 1e-3 of the fluid count).  Where the spreads vanish (no
  fluid left, or fluid at rest with C = 0) the floors are the bound.
"""
import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes

pytestmark = pytest.mark.gpu


def _ctx(wl):
    p = vtp.params_from_json(wl.solver, wl.config_json())
    ctx = vtp.Context(p)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, o.vol_frac, o.scene)
    ctx.initialize()
    return p, ctx


def test_dfsph_cfg1_cfl_1000_ticks_every_solve_converges():
    wl = scenes.cfg1_dfsph(seed=0, cfl=True)
    p, ctx = _ctx(wl)
    print(f"DFSPH 64^3 CFL: {wl.n} particles ({wl.n_fluid} fluid), 1000 ticks", flush=True)
    fluid = ctx.read("mat") == vtp.MAT_FLUID
    worst = {"div": 0, "inc": 0}
    for t in range(1000):
        ctx.step(1)
        s = ctx.stats()
        assert s["crash_flags"] == 0, f"tick {t + 1}: crash flags {s['crash_flags']:#x}"
        assert s["dead_particles"] == 0, f"tick {t + 1}: {s['dead_particles']} particles off the grid"
        assert s["div_iterations"] <= 101 and s["incomp_iterations"] <= 101
        assert s["div_residual"] < p.div_free_threshold, (t + 1, s["div_residual"])
        assert s["incomp_residual"] < p.incomp_threshold, (t + 1, s["incomp_residual"])
        worst["div"] = max(worst["div"], s["div_iterations"])
        worst["inc"] = max(worst["inc"], s["incomp_iterations"])
        if (t + 1) % 100 == 0:
            cr = ctx.read("compression_ratio")[fluid]
            bulk = cr[cr >= 0.9]
            m, q = float(bulk.mean()), float(np.quantile(bulk, 0.999))
            print(f"tick {t + 1}: t = {s['sim_time']:.4f} s, dt = {s['dt']:.2e}, iterations (div, density) "
                  f"{s['div_iterations']}, {s['incomp_iterations']}; bulk cr mean {m:.5f}, 99.9% {q:.4f}",
                  flush=True)
            assert abs(m - 1.0) <= 0.01, (t + 1, m)
            assert q <= 1.05, (t + 1, q)
    assert s["sim_time"] > 0.5   # the CFL step moved time on (dt <= dtMax = 0.002)
    print(f"most iterations in one solve: div {worst['div']}, density {worst['inc']}", flush=True)
    ctx.close()


def _pbf_stats(pos, dens, mass, C, fluid, lb, ub, rho0):
    """Fluid on the grid: its count, mean density / rho0, mean and 99th
    percentile of C, and SPH volume (sum of m / rho)."""
    live = fluid & np.all((pos >= lb) & (pos < ub), axis=1)
    if not live.any():
        return np.array([0.0, 1.0, 0.0, 0.0, 0.0])
    d, m, c = dens[live].astype(np.float64), mass[live].astype(np.float64), C[live].astype(np.float64)
    return np.array([live.sum(), d.mean() / rho0, c.mean(), np.quantile(c, 0.99), (m / d).sum()])


PBF_STATS = ("fluid on the grid", "mean density / rho0", "mean C", "99th percentile of C", "SPH volume")
# floors of the bounds: 5e-5 of rest density for the density and the volume
# (relative), 5e-5 for the mean of C and 2e-3 for its 99th percentile, and
# 1e-3 of the fluid count
PBF_FLOOR_REL = np.array([1e-3, 5e-5, 0.0, 0.0, 5e-5])
PBF_FLOOR_ABS = np.array([0.0, 0.0, 5e-5, 2e-3, 0.0])


@pytest.mark.parametrize("physics", ["reference", "mass_weighted"])
def test_pbf_cfg0_500_ticks_statistics_match_oracle(physics):
    wl = scenes.cfg0_pbf(seed=0)
    if physics == "mass_weighted":
        wl.config["PBF"].setdefault("Extensions", {})["massWeighted"] = True
    p, ctx = _ctx(wl)
    rho0 = float(p.f_rest_density)
    lb = np.array(list(p.sim_space_lb), np.float64)
    ub = lb + np.array(list(p.sim_space_size), np.float64)
    print(f"PBF 32^3 ({physics}): {wl.n} particles ({wl.n_fluid} fluid), 500 ticks", flush=True)
    fluid = ctx.read("mat") == vtp.MAT_FLUID
    runs = []
    for seed in (None, 1, 2, 3):   # the FP64 oracle and three starts moved by 1e-7 m
        o = pyoracle.Oracle.from_workload(p, wl)
        o.initialize()
        if seed:
            x = o.read(0)
            x[fluid] += np.random.default_rng(seed).choice([-1e-7, 1e-7], size=(int(fluid.sum()), 3))
            o.write(0, x)
        runs.append(o)
    ids = vtp._abi.FIELDS
    fid = [ids[f][0] for f in ("pos", "density", "mass", "error")]
    samples = []   # every 10 ticks: [gpu, oracle, perturbed x3] x statistics
    for t in range(10, 501, 10):
        ctx.step(10)
        row = [_pbf_stats(ctx.read("pos"), ctx.read("density"), ctx.read("mass"), ctx.read("error"), fluid, lb, ub,
                          rho0)]
        for o in runs:
            o.step(10)
            row.append(_pbf_stats(*(o.read(f) for f in fid), fluid, lb, ub, rho0))
        samples.append(row)
        if t % 50 == 0:
            w = np.array(samples[-5:]).mean(0)   # the window's means
            g, r, q = w[0], w[1], w[2:]
            spread = np.abs(q - r).max(0)
            bound = 3 * spread + PBF_FLOOR_REL * np.abs(r) + PBF_FLOOR_ABS
            print(f"ticks {t - 49}-{t}: " + "; ".join(
                f"{PBF_STATS[k]} gpu {g[k]:.6g} oracle {r[k]:.6g} spread {spread[k]:.2e}" for k in range(5)),
                flush=True)
            bad = [PBF_STATS[k] for k in range(5) if not abs(g[k] - r[k]) <= bound[k]]
            assert not bad, (t, bad, g, r, spread)
    st = ctx.stats()
    print(f"off the grid after 500 ticks: {st['dead_particles']} particles", flush=True)
    # particles that sink through the floor leave the grid: not a crash
    # (VTP_CRASH_OUT_OF_GRID), and no other crash bit may be set
    assert st["crash_flags"] & ~vtp._abi.CRASH_OUT_OF_GRID == 0, st["crash_flags"]
    ctx.close()


def test_immct_1m_300_ticks_solves_converge_and_phases_are_conserved():
    """IMM-CT three-phase mixture at its full size and length: 100^3 fluid
    at spacing 0.01 (+591,065 wall particles), 300 ticks at dt 0.001, with
    the mass-weighted polymer stress of the benchmark (the reference's
    mass-free term blows up at this resolution within two ticks).  Every
    tick: no crash bit, nothing off the grid, both solves stop below their
    thresholds within the iteration cap (IMMCTSolver.cpp:408-492 with the
    DFSPH solve loops).  Every 50 ticks: each phase's total volume (FP32
    fractions summed in FP64) within 1e-4 of its start, the fractions of
    every particle a partition of one within 1e-5 and none below -1e-6, and
    the fluid not compressed: the 99.9th percentile of its compression ratio
    under 1.05.  The density solve corrects compression only (delta =
    max(cr - 1, 0)), so the ratio may fall below 1 where the falling slabs
    spread (bulk mean 0.982 at tick 50); it is printed, not bounded."""
    wl = scenes.cfg2_immct(seed=0)
    wl.config["IMMCT"]["Extensions"]["massWeighted"] = True
    p, ctx = _ctx(wl)
    nf = wl.n_fluid
    print(f"IMM-CT 1M: {wl.n} particles ({nf} fluid), 300 ticks", flush=True)
    s0 = ctx.read("vol_frac")[:nf].astype(np.float64).sum(0)
    worst = {"div": 0, "inc": 0, "phase": 0}
    for t in range(300):
        ctx.step(1)
        s = ctx.stats()
        assert s["crash_flags"] == 0, f"tick {t + 1}: crash flags {s['crash_flags']:#x}"
        assert s["dead_particles"] == 0, (t + 1, s["dead_particles"])
        assert s["div_iterations"] <= 101 and s["incomp_iterations"] <= 101
        assert s["div_residual"] < p.div_free_threshold, (t + 1, s["div_residual"])
        assert s["incomp_residual"] < p.incomp_threshold, (t + 1, s["incomp_residual"])
        worst["div"] = max(worst["div"], s["div_iterations"])
        worst["inc"] = max(worst["inc"], s["incomp_iterations"])
        worst["phase"] = max(worst["phase"], s["phase_iterations"])
        if (t + 1) % 50 == 0:
            f = ctx.read("vol_frac")[:nf].astype(np.float64)
            drift = np.abs(f.sum(0) / s0 - 1.0)
            part = float(np.abs(f.sum(1) - 1.0).max())
            cr = ctx.read("compression_ratio")[:nf].astype(np.float64)
            bulk = cr[cr >= 0.9]
            q = float(np.quantile(cr, 0.999))
            print(f"tick {t + 1}: phase volume drift {drift.max():.2e}, partition error {part:.2e}, "
                  f"min fraction {f.min():.2e}, bulk cr mean {bulk.mean():.5f}, 99.9% {q:.4f}; iterations (div, density, "
                  f"phase) {s['div_iterations']}, {s['incomp_iterations']}, {s['phase_iterations']}", flush=True)
            assert drift.max() <= 1e-4, (t + 1, drift)
            assert part <= 1e-5 and f.min() >= -1e-6, (t + 1, part, f.min())
            assert q <= 1.05, (t + 1, q)
    print(f"most iterations in one tick: div {worst['div']}, density {worst['inc']}, phase {worst['phase']}",
          flush=True)
    ctx.close()
