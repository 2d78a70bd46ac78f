"""The C++ host API (Solver / Object / ObjectManager / ModelHandler /
ExportUtil over the C ABI), host-only parts.

The generators are checked bit-for-bit against the NumPy restatement in
vtp.scenes (itself pinned to the particle counts of the reference's
ModelHandler rules), the PLY reader against files in all three PLY encodings,
and the templates against the reference's template values.
"""
import json
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "vt-physics_amd" / "bin"
RUNNER = BIN / "scene_runner"

pytestmark = pytest.mark.skipif(not RUNNER.exists(), reason="host examples not built (make -C vt-physics_amd)")


def _generate(tmp_path, objects):
    scene = {"OBJECTS": objects}
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(scene))
    out = tmp_path / "gen"
    out.mkdir()
    r = subprocess.run([str(RUNNER), str(p), "--generate-only", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    counts = json.loads(r.stdout.strip().splitlines()[-1])["objects"]
    arrays = [np.fromfile(out / f"{i}.bin", dtype=np.float32).reshape(-1, 3) for i in range(len(objects))]
    assert [a.shape[0] for a in arrays] == counts
    return arrays


def _obj(t, comp, mat=0):
    cfg = dict(comp)
    cfg.setdefault("epmMaterial", mat)
    return {"objTransferType": t, "ObjectComponentConfig": cfg,
            "SolverObjectComponentConfig": {"solverType": 0, "exportFlag": False, "velocityStart": [0, 0, 0],
                                            "colorStart": [0, 1, 0]}}


GEOMS = [
    (2, {"particleRadius": 0.01, "lb": [0.04, 0.04, 0.04], "size": [0.64, 0.64, 0.64]}),
    (5, {"particleRadius": 0.01, "lb": [0, 0, 0], "size": [1, 1, 1], "layerNum": 2}),
    (5, {"particleRadius": 0.005, "lb": [0, 0, 0], "size": [2, 1, 1], "layerNum": 2}),
    (1, {"particleRadius": 0.02, "volumeCenter": [0.3, 0.5, -0.1], "volumeRadius": 0.37}),
    (3, {"particleRadius": 0.015, "center": [0.1, 0.2, 0.3], "bottomAreaRadius": 0.3, "height": 0.5}),
    (4, {"particleRadius": 0.01, "lb": [-0.5, 0.1, -0.25], "size": [1.0, 0.0, 0.5], "layerNum": 3}),
    (2, {"particleRadius": 0.05, "lb": [-1, -1, -1], "size": [2, 2, 2]}),
]


def test_generators_match_numpy_restatement(tmp_path):
    from vtp import scenes
    arrays = _generate(tmp_path, [_obj(t, c) for t, c in GEOMS])
    for (t, c), got in zip(GEOMS, arrays):
        cfg = dict(c, genType=t)
        want = scenes.generate(cfg)
        assert got.shape == want.shape, (t, got.shape, want.shape)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), t


def test_reference_lattice_counts(tmp_path):
    """Config-0 fluid block (32^3) and its [0,1]^3 two-layer box (35,315)."""
    arrays = _generate(tmp_path, [_obj(*GEOMS[0]), _obj(*GEOMS[1], mat=1), _obj(*GEOMS[2], mat=1)])
    assert arrays[0].shape[0] == 32768
    assert arrays[1].shape[0] == 35315
    assert arrays[2].shape[0] == 244065


def _write_ply(path, pts, fmt, extra_face=False):
    n = pts.shape[0]
    head = ["ply", f"format {fmt} 1.0", "comment test", f"element vertex {n}", "property double x",
            "property float y", "property float z", "property uchar red"]
    if extra_face:
        head += ["element face 1", "property list uchar int vertex_indices"]
    head.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for p in pts:
                f.write(f"{float(p[0])!r} {float(p[1])!r} {float(p[2])!r} 7\n".encode())
            if extra_face:
                f.write(b"3 0 1 2\n")
        else:
            e = "<" if fmt == "binary_little_endian" else ">"
            for p in pts:
                f.write(struct.pack(e + "dffB", float(p[0]), float(p[1]), float(p[2]), 7))
            if extra_face:
                f.write(struct.pack(e + "Biii", 3, 0, 1, 2))


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_particle_import(tmp_path, fmt):
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1, 1, size=(257, 3)).astype(np.float32)
    ply = tmp_path / "cloud.PLY"
    _write_ply(ply, pts, fmt, extra_face=True)
    (got,) = _generate(tmp_path, [_obj(0, {"particleRadius": 0.01, "particleGeometryPath": str(ply)})])
    assert np.array_equal(got, pts)


def test_templates_are_the_reference_defaults():
    import vtp
    pbf = vtp.config_template("PBF")
    assert pbf["PBF"]["Required"]["iterationNum"] == 10
    assert pbf["PBF"]["Required"]["XSPH_k"] == pytest.approx(0.02)
    assert pbf["EXPORT"]["SolverRequired"]["exportGroupPolicy"] == "MERGE"
    immct = vtp.config_template("IMMCT")
    assert immct["IMMCT"]["Required"]["phaseRestDensity"] == [900, 1000]
    for i, name in enumerate(["PBF", "DFSPH", "IMM", "IMMCT"]):
        oc = vtp.object_config_template(name)
        assert oc["solverType"] == i
        assert set(oc) >= {"solverType", "exportFlag", "velocityStart"}
    assert vtp.object_config_template("IMM")["phaseFraction"] == [0.1, 0.1]


def test_solver_rejects_incomplete_config(tmp_path):
    """setConfig fails on a missing Required key; run is not reached."""
    scene = {"PBF": {"Required": {"animationTime": 1.0}}, "EXPORT": {},
             "OBJECTS": [_obj(*GEOMS[6])]}
    p = tmp_path / "bad.json"
    p.write_text(json.dumps(scene))
    r = subprocess.run([str(RUNNER), str(p), "--steps", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1
    assert "missing" in r.stderr and "timeStep" in r.stderr
    assert '"ok": false' in r.stdout
