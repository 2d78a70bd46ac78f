"""The measurement side of bench.py, on the host: the per-kernel byte split
against the per-tick formula it is taken from, the rocprofv3 kernel-trace and
counter summaries on small hand-made CSV files (the columns rocprofv3 writes),
and the one-GPU guard of the command line."""
import csv
import io
import json
import subprocess
import sys
from contextlib import redirect_stdout
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


@pytest.mark.parametrize("n,nf", [(1_321_440, 1_000_000), (7_809_555, 4_903_495), (1000, 0)])
def test_pass_bytes_split_the_per_iteration_term(n, nf):
    """One constraint iteration adds 32 N + 16 N_f to the tick; the density +
    lambda pass and the correction pass share exactly that, and the list
    build with the folded first lambda pass is credited with the lambda share."""
    per_iter = bench.pbf_step_bytes(n, nf, 5000, 4, 1) - bench.pbf_step_bytes(n, nf, 5000, 3, 1)
    assert per_iter == 32 * n + 16 * nf
    assert bench.kernel_bytes("pbf_lambda", n, nf) + bench.kernel_bytes("pbf_dx", n, nf) == per_iter
    assert bench.kernel_bytes("ugns_neighbors", n, nf, fold=True) == bench.kernel_bytes("pbf_lambda", n, nf)
    assert bench.kernel_bytes("ugns_neighbors", n, nf) is None
    # XSPH + apply is the only other per-slot term of the tick: 28 N with no fluid
    assert bench.pbf_step_bytes(n, 0, 0, 0, 0) == bench.kernel_bytes("pbf_xsph", n, 0) == 28 * n
    # a kernel's own fields are never fewer bytes than its share of the formula
    for k in ("pbf_lambda", "pbf_dx", "pbf_xsph"):
        assert bench.kernel_bytes_own_fields(k, n, nf) >= bench.kernel_bytes(k, n, nf)


def _write_csv(path, fields, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(fields, r)))


def test_trace_summary_assigns_each_leg_from_its_context_initialisation(tmp_path):
    """A leg starts at the k_iota that precedes its first own kernel, so the
    search of its first tick (here k_cell_ranges) counts to it and not to the
    leg before."""
    seq = [
        ("k_iota", 10), ("k_build_neighbors<true, 1u>", 200), ("vtp::k_pbf_lambda<1u>", 50),
        ("vtp::k_pbf_dx<1u>", 50),                                   # headline
        ("k_iota", 11), ("k_cell_ranges", 40), ("vtp::k_df_precompute", 100),
        ("vtp::k_df_delta_kappa<1u>", 90),                            # DFSPH
        ("k_iota", 12), ("vtp::k_pbf_lambda<1u>", 400),               # batched
        ("k_iota", 13), ("vtp::k_pbf_dx<1u>", 30), ("vtp::k_adj_dx_rev", 60),   # adjoint
        ("k_iota", 14), ("vtp::k_imm_fluxes<3>", 300),                # IMM-CT
    ]
    rows, t = [], 1000
    for name, dur in seq:
        rows.append((f"void {name}(float4 const*, unsigned int)", t, t + dur * 1000, 256))
        t += dur * 1000 + 500
    trace = tmp_path / "kernel_trace.csv"
    _write_csv(trace, ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X"], rows)
    buf = io.StringIO()
    with redirect_stdout(buf):
        bench.trace_summary(str(trace))
    legs, cur = {}, None
    for line in buf.getvalue().splitlines():
        if line.startswith("### "):
            cur = line[4:].split(":")[0]
            legs[cur] = []
        elif line.startswith("| `"):
            legs[cur].append(line.split("`")[1])
    assert list(legs) == ["pbf_1m (headline)", "dfsph_1m", "batched_64", "adjoint", "immct_1m"]
    assert sorted(legs["pbf_1m (headline)"]) == sorted(["k_build_neighbors<true, 1u>", "k_pbf_lambda<1u>",
                                                        "k_pbf_dx<1u>", "k_iota"])
    assert sorted(legs["dfsph_1m"]) == sorted(["k_iota", "k_cell_ranges", "k_df_precompute",
                                               "k_df_delta_kappa<1u>"])
    assert sorted(legs["batched_64"]) == ["k_iota", "k_pbf_lambda<1u>"]
    assert sorted(legs["adjoint"]) == ["k_adj_dx_rev", "k_iota", "k_pbf_dx<1u>"]
    assert sorted(legs["immct_1m"]) == ["k_imm_fluxes<3>", "k_iota"]
    # leg totals: each k_iota counts to the leg it opens (10 + 200 + 50 + 50 us, 11 + 40 + 100 + 90 us)
    assert "### pbf_1m (headline): 0.310 ms" in buf.getvalue()
    assert "### dfsph_1m: 0.241 ms" in buf.getvalue()


def test_traffic_summary_scales_the_per_tick_median_by_the_copy_reading(tmp_path):
    """Per kernel: the median KiB per dispatch of every instantiation launched
    at least half as often as the most launched one (the build and its 32-bit
    fallback, not the search at initialisation), summed, times
    (copy bytes / what the counter read for the copy)."""
    fields = ["Kernel_Name", "Counter_Name", "Counter_Value"]

    def rows(counter, table):
        return [(k, counter, v) for k, vals in table.items() for v in vals]

    fetch = {"k_build_neighbors<true, 1u>(...)": [100, 130, 110, 120],     # median 120
             "k_build_neighbors<false, 1u>(...)": [12, 10],                # launched 2 of 4: counted
             "k_build_neighbors<true, 0u>(...)": [999],                    # initialisation only
             "k_pbf_lambda<1u>(...)": [40, 44, 42],
             "k_pbf_dx<1u>(...)": [50, 50], "k_pbf_xsph(...)": [70]}
    write = {"k_build_neighbors<true, 1u>(...)": [30, 30, 30, 30],
             "k_build_neighbors<false, 1u>(...)": [4, 4],
             "k_pbf_lambda<1u>(...)": [8, 8, 8],
             "k_pbf_dx<1u>(...)": [20, 20], "k_pbf_xsph(...)": [25]}
    copy_kib = (1 << 30) / 1024
    paths = {k: tmp_path / f"{k}.csv" for k in ("f", "w", "cf", "cw")}
    _write_csv(paths["f"], fields, rows("FETCH_SIZE", fetch))
    _write_csv(paths["w"], fields, rows("WRITE_SIZE", write))
    _write_csv(paths["cf"], fields, [("copy_kernel", "FETCH_SIZE", 2 * copy_kib)] * 3)   # reads twice
    _write_csv(paths["cw"], fields, [("copy_kernel", "WRITE_SIZE", copy_kib)] * 3)
    out = tmp_path / "traffic.json"
    committed = bench.TRAFFIC_JSON.read_bytes()
    with redirect_stdout(io.StringIO()):
        bench.traffic_summary(str(paths["f"]), str(paths["w"]), str(paths["cf"]), str(paths["cw"]),
                              workload="pbf", out=out)
    doc = json.loads(out.read_text())
    assert doc["workload"] == "pbf_dam_break_1m"
    assert doc["fetch_scale"] == 0.5 and doc["write_scale"] == 1.0
    k = doc["kernels"]
    assert k["ugns_neighbors"]["read_bytes"] == (120 + 12) * 1024 // 2
    assert k["ugns_neighbors"]["write_bytes"] == (30 + 4) * 1024
    assert k["pbf_lambda"] == {"read_bytes": 42 * 512, "write_bytes": 8 * 1024, "bytes": 42 * 512 + 8 * 1024}
    assert k["pbf_dx"]["read_bytes"] == 50 * 512 and k["pbf_xsph"]["write_bytes"] == 25 * 1024
    # the committed summary is not touched by a run that names its own file
    assert bench.TRAFFIC_JSON.read_bytes() == committed


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
