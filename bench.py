#!/usr/bin/env python3
"""Benchmark of the particle-fluid step on one MI355X.

Headline: PBF dam break at 1,000,000 fluid particles (100^3 lattice at
spacing 0.02, 321,440 wall particles, config-0 physics: dt 0.004, 4
constraint iterations, XSPH 0.01), inputs resident in HBM, the tick replayed
as a hipGraph.  value = fluid particle-steps per second.

Also reported: the dominant kernel's achieved algorithmic bandwidth against
the 8 TB/s HBM peak (hipEvents on the context's stream), the FP64 oracle
timed on the host cores over a bounded sample of the same workload, and the
other parts of the metric: DFSPH at 1M fluid particles, the batched
64-scene gain (one context holding all scenes vs. one context per scene),
the adjoint cost ((forward + backward) / forward of the checkpointed
100-tick rollout), and the three-phase IMM-CT mixture at 1M fluid particles.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "vt-physics_amd"))

HBM_PEAK_GBS = 8000.0


def pbf_step_bytes(n, nf, cells, K, passes):
    """Algorithmic HBM bytes of one PBF tick (compulsory SoA traffic only)."""
    return nf * (48 + 88 + 16 * passes) + 8 * cells + K * (32 * n + 16 * nf) + 28 * n + 24 * nf


def kernel_bytes(name, n, nf, fold=False):
    """Algorithmic bytes per launch of the PBF neighbour passes, split from the
    per-tick formula above: an iteration's 32 N + 16 N_f is the density+lambda
    pass (x, m in -> lambda out: 16 N + 4 N_f) plus the correction pass
    (x, lambda in -> x out: 16 N + 12 N_f); XSPH + apply is 28 N + 24 N_f.
    Neighbour re-reads and lists are not credited, so the list build has no
    share of its own; with `fold` (option build_lambda) it is credited with
    the first iteration's density+lambda pass it computes."""
    if name == "pbf_lambda":
        return 16 * n + 4 * nf
    if name == "pbf_dx":
        return 16 * n + 12 * nf
    if name == "pbf_xsph":
        return 28 * n + 24 * nf
    if name == "ugns_neighbors" and fold:
        return 16 * n + 4 * nf
    return None


def kernel_bytes_own_fields(name, n, nf, fold=False):
    """Bytes per launch counting every field the kernel itself reads and
    writes once (material words, both halves of the (x, lambda) records,
    velocities), the per-kernel count of the design notes."""
    if name == "pbf_lambda":      # x,m + mat in; lambda out (all); rho, C out (fluid)
        return 24 * n + 8 * nf
    if name == "pbf_dx":          # x,m + lambda + mat in (all); x out, v in/out (fluid)
        return 24 * n + 48 * nf
    if name == "pbf_xsph":        # x, v, mat in (all); x, v out (all)
        return 68 * n
    if name == "ugns_neighbors":  # x + key in; count word out; with the tick's first lambda pass
        # mat in, lambda out (all), rho and C out (fluid)
        return 36 * n + 8 * nf if fold else 24 * n
    return None


def dfsph_step_bytes(n, nf, cells, passes, it_div, it_inc):
    """Algorithmic HBM bytes of one DFSPH tick: grid, the alpha + compression
    ratio pass, per solver iteration the delta+kappa and v* passes, two
    pressure applications, the non-pressure forces and the position update."""
    return (nf * (88 + 16 * passes) + 8 * cells + 20 * n + 8 * nf + (it_div + it_inc) * (48 * n + 40 * nf)
            + 2 * 48 * nf + 150 * nf + 36 * nf)


def immct_step_bytes(n, nf, cells, passes, it_div, it_inc, P, phase_passes, ct_passes=2):
    """The DFSPH tick's bytes plus P (4 + 4 + 4 + 12 + 12 + 12) bytes per
    fluid particle for every pass that touches the phases (fractions in and
    out, fluxes, phase velocities, drift and phase acceleration) and
    3 x 36 bytes per fluid particle (Q, L, stress) for each of the two
    conformation-tensor passes."""
    return (dfsph_step_bytes(n, nf, cells, passes, it_div, it_inc) + phase_passes * P * 48 * nf
            + ct_passes * 3 * 36 * nf)


def build_ctx(vtp, wl, graph=True):
    params = vtp.params_from_json(wl.solver, wl.config_json())
    ctx = vtp.Context(params)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, o.vol_frac, o.scene)
    ctx.set_option("graph", 1 if graph else 0)
    # workgroup size of every neighbour pass timed at initialize (results
    # bit-identical to any other choice)
    ctx.set_option("autotune", 1)
    ctx.initialize()
    return ctx, params


def timed_steps(torch, ctx, steps):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def stage_breakdown(ctx, steps, image=None):
    """Per-stage device ms per tick from hipEvents on the context's stream
    (eager launches).  With `image` (the state image taken at the start of
    the timed ticks) the same ticks are replayed, so the split covers the
    timed work (a resume is bit-exact, iteration counts included)."""
    if image is not None:
        ctx.load_state(image)
    ctx.set_option("graph", 0)
    ctx.set_option("stage_timing", 1)
    ctx.reset_stage_times()
    ctx.step(steps)
    st = ctx.stage_times()
    ctx.set_option("stage_timing", 0)
    return {k: round(v[0] / steps, 4) for k, v in st.items()}


def key_passes(cells):
    return 3 if cells >= (1 << 16) else 2


def bench_dfsph(torch, vtp, scenes, steps, warmup):
    wl = scenes.dfsph_1m(seed=0)
    ctx, params = build_ctx(vtp, wl, graph=False)
    _, cells = ctx.grid_info()
    ctx.step(warmup)
    st0 = ctx.stats()
    image = ctx.save_state()
    el = timed_steps(torch, ctx, steps)
    st = ctx.stats()
    stages = stage_breakdown(ctx, steps, image)
    shapes = ctx.launch_shapes(vtp._abi.DFSPH)
    ctx.close()
    it_div = (st["total_div_iterations"] - st0["total_div_iterations"]) / steps
    it_inc = (st["total_incomp_iterations"] - st0["total_incomp_iterations"]) / steps
    ms = el * 1e3 / steps
    sb = dfsph_step_bytes(wl.n, wl.n_fluid, cells, key_passes(cells), it_div, it_inc)
    return {"workload": wl.name, "fluid_particles": wl.n_fluid, "boundary_particles": wl.n - wl.n_fluid,
            "value": round(wl.n_fluid * steps / el, 1), "unit": "particle-steps/s",
            "ms_per_step": round(ms, 4), "mean_div_iterations": it_div,
            "mean_incomp_iterations": it_inc, "steps": steps, "warmup": warmup,
            "step_algorithmic_bytes": int(sb), "step_hbm_frac": round(sb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "dead_particles": st["dead_particles"], "stages_ms_per_step": stages, "launch_shapes": shapes}


def bench_immct(torch, vtp, scenes, steps, warmup):
    wl = scenes.cfg2_immct(seed=0)
    # the reference's polymer-stress term has no neighbour mass and blows up at
    # r = 0.005 within two ticks; the timed run uses the mass-weighted form
    wl.config["IMMCT"]["Extensions"]["massWeighted"] = True
    ctx, params = build_ctx(vtp, wl, graph=False)
    _, cells = ctx.grid_info()
    ctx.step(warmup)
    st0 = ctx.stats()
    image = ctx.save_state()
    el = timed_steps(torch, ctx, steps)
    st = ctx.stats()
    stages = stage_breakdown(ctx, steps, image)
    shapes = ctx.launch_shapes(vtp._abi.IMMCT)
    ctx.close()
    it_div = (st["total_div_iterations"] - st0["total_div_iterations"]) / steps
    it_inc = (st["total_incomp_iterations"] - st0["total_incomp_iterations"]) / steps
    ms = el * 1e3 / steps
    # phase passes: pressure -> phases (x2), viscosity, tension + phase velocity,
    # stress, phase update, and one flux pass per negative-fraction round
    phase_passes = 6 + max(1, st["phase_iterations"])
    sb = immct_step_bytes(wl.n, wl.n_fluid, cells, key_passes(cells), it_div, it_inc, 3, phase_passes)
    return {"workload": wl.name, "fluid_particles": wl.n_fluid, "boundary_particles": wl.n - wl.n_fluid,
            "stages_ms_per_step": stages, "phases": 3, "value": round(wl.n_fluid * steps / el, 1), "unit": "particle-steps/s",
            "ms_per_step": round(ms, 4), "steps": steps, "warmup": warmup,
            "mean_div_iterations": it_div, "mean_incomp_iterations": it_inc,
            "step_algorithmic_bytes": int(sb), "step_hbm_frac": round(sb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "mass_weighted_stress": True, "launch_shapes": shapes, "last_iterations": [st["div_iterations"], st["incomp_iterations"],
                                                              st["phase_iterations"]]}


def bench_batched(torch, vtp, scenes, steps, warmup):
    wl = scenes.cfg3_batched(seed=0)
    ctx, params = build_ctx(vtp, wl, graph=True)
    ctx.step(warmup)
    t_batched = timed_steps(torch, ctx, steps)
    ctx.close()
    # the same scenes, one context (one launch sequence) per scene
    per = {}
    for o in wl.objects:
        per.setdefault(o.scene, []).append(o)
    ctxs = []
    for sc, objs in sorted(per.items()):
        c = vtp.Context(params)
        for o in objs:
            c.add_particles(o.pos, o.mat, o.vel, None, 0)
        c.set_option("graph", 1)
        c.set_option("autotune", 1)
        c.initialize()
        c.step(warmup)
        ctxs.append(c)
    # each scene's ticks back to back with one synchronisation per scene (as
    # the batched run has one for all of them)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(steps)
    torch.cuda.synchronize()
    t_single = time.perf_counter() - t0
    # the same, one tick of every scene in turn (a synchronisation per scene tick)
    t0 = time.perf_counter()
    for _ in range(steps):
        for c in ctxs:
            c.step(1)
    torch.cuda.synchronize()
    t_inter = time.perf_counter() - t0
    # the per-scene ticks as plain launch sequences (no graph replay)
    for c in ctxs:
        c.set_option("graph", 0)
        c.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for c in ctxs:
        c.step(steps)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    return {"workload": wl.name, "scenes": len(per), "fluid_particles": wl.n_fluid,
            "boundary_particles": wl.n - wl.n_fluid, "steps": steps,
            "batched_ms_per_step": round(t_batched * 1e3 / steps, 4),
            "one_context_per_scene_ms_per_step": round(t_single * 1e3 / steps, 4),
            "gain": round(t_single / t_batched, 3),
            "one_context_per_scene_interleaved_ms_per_step": round(t_inter * 1e3 / steps, 4),
            "gain_interleaved": round(t_inter / t_batched, 3),
            "one_context_per_scene_eager_ms_per_step": round(t_eager * 1e3 / steps, 4),
            "gain_eager": round(t_eager / t_batched, 3),
            "batched_value": round(wl.n_fluid * steps / t_batched, 1), "unit": "particle-steps/s"}


def bench_adjoint(torch, vtp, scenes, nsteps=100, ckpt=10):
    wl = scenes.cfg4_adjoint(seed=0)
    ctx, params = build_ctx(vtp, wl, graph=True)
    x0, v0 = ctx.read("pos"), ctx.read("vel")
    ctx.rollout_fwd(10, ckpt)          # warm-up (allocations, graph capture)
    ctx.write("pos", x0)
    ctx.write("vel", v0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.rollout_fwd(nsteps, ckpt)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    x = ctx.read("pos")
    nf = wl.n_fluid
    import numpy as np
    gx = np.zeros_like(x)
    target = np.array([0.7, 0.2, 0.5], np.float32)
    gx[:nf] = 2 * (x[:nf].mean(0) - target) / nf
    loss_d, gl = ctx.density_loss()
    ctx.rollout_bwd(gx + gl, None)      # warm-up of the tape allocation
    ctx.write("pos", x0)
    ctx.write("vel", v0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.rollout_fwd(nsteps, ckpt)
    _, gl = ctx.density_loss()
    gv0, _, gk = ctx.rollout_bwd(gx + gl, None)
    torch.cuda.synchronize()
    t_both = time.perf_counter() - t0
    ctx.close()
    return {"workload": wl.name, "fluid_particles": nf, "steps": nsteps, "checkpoint_every": ckpt,
            "max_neighbor_num": params.max_neighbor_num,
            "forward_ms": round(t_fwd * 1e3, 3), "forward_backward_ms": round(t_both * 1e3, 3),
            "cost_ratio": round(t_both / t_fwd, 3), "grad_xsph_k": gk}


def cpu_baseline(wl, params, budget_s=20.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    threads = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():   # the host share this job may use
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    pyoracle.set_threads(threads)
    orc = pyoracle.Oracle.from_workload(params, wl)
    orc.initialize()
    orc.step(1)  # first tick outside the timing (allocations)
    steps = 0
    t0 = time.perf_counter()
    while True:
        orc.step(1)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 50:
            break
    return {"value": wl.n_fluid * steps / el, "unit": "particle-steps/s", "cores": pyoracle.get_threads(),
            "kind": "port", "cpu_model": cpu_model(), "host_logical_cpus": os.cpu_count(),
            "sample": f"FP64 oracle, {wl.name}: {steps} ticks after 1 untimed tick ({el:.1f} s)"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


STAGE_KERNEL = {"pbf_lambda": "k_pbf_lambda", "pbf_dx": "k_pbf_dx", "pbf_xsph": "k_pbf_xsph",
                "ugns_neighbors": "k_build_neighbors"}
TRAFFIC_JSON = ROOT / "profiles" / "pbf_dam_break_1m_traffic.json"
# per workload: the stages whose kernels are summarised, the output file
TRAFFIC_WORKLOADS = {
    "pbf": ("pbf_dam_break_1m", STAGE_KERNEL, TRAFFIC_JSON),
    "dfsph": ("dfsph_dam_break_1m",
              {"df_precompute": "k_df_precompute", "df_delta_kappa": "k_df_delta_kappa",
               "df_update_vadv": "k_df_update_vadv", "df_vis_normal": "k_df_vis_normal",
               "ugns_neighbors": "k_build_neighbors"},
              ROOT / "profiles" / "dfsph_dam_break_1m_traffic.json"),
}


def traffic_summary(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv, workload="pbf", out=None):
    """HBM bytes per launch of the neighbour-pass kernels from two rocprofv3
    --pmc runs (FETCH_SIZE, WRITE_SIZE, KiB per dispatch; the median over a
    kernel's launches), each scaled by what the same counter reads for a
    plain device copy of known size.  Written to the workload's file under
    profiles/ unless `out` names another."""
    wl_name, stage_kernel, default_out = TRAFFIC_WORKLOADS[workload]
    out = default_out if out is None else out
    import csv
    from collections import defaultdict

    def dispatches(path, counter):
        acc = defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        return acc

    def per_kernel(path, counter):
        return {k: sum(v) / len(v) for k, v in dispatches(path, counter).items()}

    def per_tick(path, counter, kname):
        """KiB per launch, summed over the kernel's instantiations launched
        every tick (those launched at least half as often as the most
        launched one: the build and its 32-bit fallback pass, not the
        search at initialisation)."""
        acc = {k: v for k, v in dispatches(path, counter).items() if kname in k}
        if not acc:
            return None
        top = max(len(v) for v in acc.values())
        return sum(sorted(v)[len(v) // 2] for v in acc.values() if 2 * len(v) >= top)

    cf = max(per_kernel(copy_fetch_csv, "FETCH_SIZE").values())
    cw = max(per_kernel(copy_write_csv, "WRITE_SIZE").values())
    copy_kib = (1 << 30) / 1024   # calibrate_copy: 1 GiB read and 1 GiB written per dispatch
    fetch_scale, write_scale = copy_kib / cf, copy_kib / cw
    kernels = {}
    for stage, kname in stage_kernel.items():
        fk, wk = per_tick(fetch_csv, "FETCH_SIZE", kname), per_tick(write_csv, "WRITE_SIZE", kname)
        if fk is not None and wk is not None:
            rb, wb = fk * 1024 * fetch_scale, wk * 1024 * write_scale
            kernels[stage] = {"read_bytes": round(rb), "write_bytes": round(wb), "bytes": round(rb + wb)}
    doc = {"workload": wl_name, "fetch_scale": round(fetch_scale, 4), "write_scale": round(write_scale, 4),
           "kernels": kernels}
    Path(out).write_text(json.dumps(doc, indent=1) + "\n")
    print(json.dumps(doc))


LEG_FIRST_KERNEL = (("dfsph_1m", "k_df_"), ("batched_64", "k_pbf_"), ("adjoint", "k_adj_"), ("immct_1m", "k_imm"))


def trace_summary(trace_csv):
    """Markdown tables from a rocprofv3 --kernel-trace CSV of a default bench
    run: per kernel and leg, launches and microseconds per launch.  Legs run
    in the bench's order (PBF 1M headline, DFSPH 1M, batched, adjoint, IMM-CT
    1M); a leg starts at the k_iota (the first kernel of a context's
    initialisation) that precedes the first dispatch of its own kernel
    family, so its contexts' first neighbour search counts to it."""
    import csv
    from collections import defaultdict

    rows = sorted(((int(r["Start_Timestamp"]), int(r["End_Timestamp"]),
                    r["Kernel_Name"].split("(")[0].replace("void ", "").replace("vtp::", ""),
                    int(r["Grid_Size_X"]))
                   for r in csv.DictReader(open(trace_csv))))
    starts, lo = [], 0
    for leg_name, prefix in LEG_FIRST_KERNEL:
        first = next((k for k in range(lo, len(rows)) if rows[k][2].startswith(prefix)), None)
        if first is None:
            break
        init = next((k for k in range(first, lo - 1, -1) if rows[k][2] == "k_iota"), first)
        starts.append((init, leg_name))
        lo = first
    legs, leg = [], "pbf_1m (headline)"
    for k, (t0, t1, name, grid) in enumerate(rows):
        if starts and k == starts[0][0]:
            leg = starts.pop(0)[1]
        legs.append((leg, name, grid, t1 - t0))
    acc = defaultdict(list)
    for leg, name, grid, d in legs:
        acc[(leg, name, grid)].append(d)
    order = [k[0] for k in LEG_FIRST_KERNEL]
    for leg in ["pbf_1m (headline)"] + order:
        items = [(k, v) for k, v in acc.items() if k[0] == leg]
        if not items:
            continue
        tot = sum(sum(v) for _, v in items)
        print(f"\n### {leg}: {tot / 1e6:.3f} ms of kernel time\n")
        print("| kernel | grid | launches | avg us | min us | max us | total ms | % of leg |")
        print("|---|---:|---:|---:|---:|---:|---:|---:|")
        for (leg_, name, grid), v in sorted(items, key=lambda kv: -sum(kv[1]))[:16]:
            if sum(v) < 0.005 * tot:
                continue
            print(f"| `{name}` | {grid} | {len(v)} | {sum(v) / len(v) / 1e3:.2f} | {min(v) / 1e3:.2f} | "
                  f"{max(v) / 1e3:.2f} | {sum(v) / 1e6:.3f} | {100 * sum(v) / tot:.2f} |")


def calibrate_copy(torch, nbytes=1 << 30, reps=5):
    """Plain device-to-device copies of a known size: the reference reading for
    scaling the FETCH_SIZE / WRITE_SIZE counters of the kernels above."""
    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    print(json.dumps({"copy_bytes_per_dispatch": nbytes, "read_kib": nbytes / 1024, "write_kib": nbytes / 1024,
                      "dispatches": reps, "GBps": round(2 * nbytes * reps / el / 1e9, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--only-headline", action="store_true", help="skip the DFSPH / batched / adjoint / IMM-CT legs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 1 GiB device copies (counter calibration for the traffic figure)")
    ap.add_argument("--traffic-summary", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="turn rocprofv3 --pmc counter files into " + str(TRAFFIC_JSON.relative_to(ROOT)) +
                    " (or the --traffic-workload's file)")
    ap.add_argument("--traffic-workload", choices=sorted(TRAFFIC_WORKLOADS), default="pbf",
                    help="which run the counter files come from: the headline (pbf) or --only-dfsph (dfsph)")
    ap.add_argument("--only-dfsph", action="store_true",
                    help="run only the DFSPH 1M leg and print its JSON (for profiling that leg)")
    ap.add_argument("--trace-summary", metavar="KERNEL_TRACE_CSV",
                    help="print per-leg kernel tables from a rocprofv3 --kernel-trace CSV of a bench run")
    args = ap.parse_args()
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary, workload=args.traffic_workload)
        return
    if args.trace_summary:
        trace_summary(args.trace_summary)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1 "
              "(multi-GPU sharding is a design note, not part of this build)", file=sys.stderr)
        sys.exit(2)

    import torch
    import vtp
    from vtp import scenes

    assert torch.cuda.is_available(), "bench needs an MI355X"
    torch.cuda.set_device(0)
    if args.calibrate_copy:
        calibrate_copy(torch)
        return
    if args.only_dfsph:
        print(json.dumps({"dfsph_1m": bench_dfsph(torch, vtp, scenes, args.steps, args.warmup)}))
        return
    wl = scenes.pbf_1m(seed=0)
    ctx, params = build_ctx(vtp, wl, graph=True)
    n, nf = wl.n, wl.n_fluid
    grid, cells = ctx.grid_info()
    ctx.step(args.warmup)
    torch.cuda.synchronize()
    image = ctx.save_state()          # the state entering the timed ticks
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = el * 1e3 / args.steps
    value = nf * args.steps / el

    # per-kernel device time with hipEvents on the context's stream, over the
    # same ticks: the state is rewound to the start of the timed region (a
    # resume is bit-exact) and the ticks are replayed with eager launches
    ctx.load_state(image)
    ctx.set_option("graph", 0)
    ctx.set_option("stage_timing", 1)
    ctx.reset_stage_times()
    ctx.step(args.steps)
    stages = ctx.stage_times()
    ctx.set_option("stage_timing", 0)
    total = sum(ms for ms, _ in stages.values())
    dom = max(stages, key=lambda k: stages[k][0])
    dom_ms = stages[dom][0] / stages[dom][1]
    fold = bool(ctx.get_option("build_lambda"))
    kb = kernel_bytes(dom, n, nf, fold)
    own = kernel_bytes_own_fields(dom, n, nf, fold)
    roof = None
    if own is not None:
        # a kernel with no share in the per-tick formula (the list build
        # without the folded lambda pass) is rated on its own fields
        basis = kb if kb is not None else own
        ach = basis / (dom_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": None, "kernel": dom,
                "kernel_ms": round(dom_ms, 5), "algorithmic_bytes": kb if kb is not None else 0,
                "bytes_basis": "per-tick formula" if kb is not None else "own fields",
                "own_field_bytes": own, "own_field_frac": round(own / (dom_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
        if TRAFFIC_JSON.exists():   # counter-measured HBM bytes per launch (rocprofv3 --pmc, calibrated)
            t = json.loads(TRAFFIC_JSON.read_text())["kernels"].get(dom)
            if t:
                roof["traffic"] = t["bytes"]
                roof["traffic_GBps"] = round(t["bytes"] / (dom_ms * 1e-3) / 1e9, 1)
                roof["traffic_source"] = str(TRAFFIC_JSON.relative_to(ROOT))
    # the same figures for every neighbour kernel of the tick (us per launch,
    # algorithmic and counter-measured fraction of the HBM peak)
    traffic = json.loads(TRAFFIC_JSON.read_text())["kernels"] if TRAFFIC_JSON.exists() else {}
    roof_all = {}
    for name in STAGE_KERNEL:
        if name not in stages or not stages[name][1]:
            continue
        ms = stages[name][0] / stages[name][1]
        ab = kernel_bytes(name, n, nf, fold)
        e = {"us_per_launch": round(ms * 1e3, 2), "launches_per_tick": round(stages[name][1] / args.steps, 2),
             "frac": round(ab / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5) if ab is not None else None,
             "own_field_frac": round(kernel_bytes_own_fields(name, n, nf, fold) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
        if name in traffic:
            e["traffic_frac"] = round(traffic[name]["bytes"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
        roof_all[name] = e
    passes = key_passes(cells * wl.scenes())
    step_bytes = pbf_step_bytes(n, nf, cells * wl.scenes(), params.iteration_num, passes)
    shapes = ctx.launch_shapes(vtp._abi.PBF)   # the tuner's choice
    st = ctx.stats()
    import numpy as np
    # ticks build fluid lists only; one more tick with every list for the statistics
    ctx.set_option("static_lists", 1)
    ctx.step(1)
    num = ctx.debug_neighbors()["neighbor_num"].astype(np.int64)   # self included, reference layout
    nbr_counts = {"mean": round(float(num[num > 0].mean()), 2), "max": int(num.max()),
                  "at_cap_fraction": round(float((num >= params.max_neighbor_num).mean()), 6)}
    ctx.close()

    extra = {}
    if not args.only_headline:
        for name, fn in (("dfsph_1m", lambda: bench_dfsph(torch, vtp, scenes, 20, 5)),
                         ("batched_64", lambda: bench_batched(torch, vtp, scenes, 10, 3)),
                         ("adjoint", lambda: bench_adjoint(torch, vtp, scenes)),
                         ("immct_1m", lambda: bench_immct(torch, vtp, scenes, 5, 2))):
            try:
                extra[name] = fn()
            except Exception as e:   # a failing leg is reported, not hidden
                extra[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"bench: {name} done", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(wl, params, args.cpu_budget)

    line = {
        "metric": "particle-steps/s (PBF, DFSPH @1M) + % HBM peak; batched gain; adjoint/fwd cost",
        "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl.name, "fluid_particles": nf, "boundary_particles": n - nf,
                   "grid": list(grid), "iterations": params.iteration_num, "dt": params.time_step,
                   "graph": True, "build_lambda": fold, "launch_shapes": shapes},
        "roofline": roof,
        "roofline_by_kernel": roof_all,
        "cpu_baseline": cpu,
        "step_hbm_frac": round(step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "stages_ms_per_step": {k: round(v[0] / args.steps, 5) for k, v in stages.items()},
        "stage_total_ms_per_step": round(total / args.steps, 4),
        "dead_particles": st["dead_particles"],
        "neighbour_counts": nbr_counts,
        "dfsph_1m": extra.get("dfsph_1m"),
        "batched_64": extra.get("batched_64"),
        "adjoint": extra.get("adjoint"),
        "immct_1m": extra.get("immct_1m"),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
