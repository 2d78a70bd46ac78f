"""The reference's own config files (its only fixtures), committed as data
under tests/golden/reference_configs/ by tests/golden/copy_reference_configs.py:

* the solver and object-component templates served by vtp_config_template /
  vtp_object_config_template equal ConfigTemplates/<Solver>/*.json
  (JsonHandler::load<Solver>ConfigTemplateJson, JSONHandler.cpp:65-71);
* each Examples/<Solver>/Scene-1/Configuration.json parses through
  vtp_params_from_json (Solver::setConfig + checkConfig) into the values the
  file holds, with the reference's derived rules (gravity -9.8 while
  Optional.enable is false, PBFSolver.cpp:53-58);
* the objects of every Scene-1 generate identically through the C++
  ObjectManager (scene_runner --generate-only) and the NumPy restatement.
Host only; the GPU runs of these scenes are in test_gpu_reference_scenes.py.
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import vtp
from vtp import scenes

ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "tests" / "golden" / "reference_configs"
RUNNER = ROOT / "vt-physics_amd" / "bin" / "scene_runner"
SOLVERS = {"PBF": "PBFSolver", "DFSPH": "DFSPHSolver", "IMM": "IMMSolver", "IMMCT": "IMMCTSolver"}


def _load(rel):
    return json.loads((REF / rel).read_text())


@pytest.mark.parametrize("solver", sorted(SOLVERS))
def test_solver_template_equals_reference_file(solver):
    d = SOLVERS[solver]
    assert vtp.config_template(solver) == _load(f"ConfigTemplates/{d}/{d}ConfigTemplate.json")


@pytest.mark.parametrize("solver", sorted(SOLVERS))
def test_object_component_template_equals_reference_file(solver):
    d = SOLVERS[solver]
    txt = vtp.load_library().vtp_object_config_template(vtp._abi.SOLVER_IDS[solver])
    assert json.loads(txt.decode()) == _load(f"ConfigTemplates/{d}/{d}ObjectComponentConfigTemplate.json")


@pytest.mark.parametrize("solver", sorted(SOLVERS))
def test_solver_template_parses_to_its_own_values(solver):
    d = SOLVERS[solver]
    doc = _load(f"ConfigTemplates/{d}/{d}ConfigTemplate.json")
    req = doc[solver]["Required"]
    p = vtp.params_from_json(solver, json.dumps(doc))
    assert p.time_step == np.float32(req["timeStep"])
    assert p.particle_radius == np.float32(req["particleRadius"])
    assert list(p.sim_space_lb) == [np.float32(v) for v in req["simSpaceLB"]]
    assert p.max_neighbor_num == req["maxNeighborNum"]


@pytest.mark.parametrize("solver", sorted(SOLVERS))
def test_example_scene_parses(solver):
    doc = _load(f"Examples/{solver}/Scene-1/Configuration.json")
    req, opt = doc[solver]["Required"], doc[solver]["Optional"]
    p = vtp.params_from_json(solver, json.dumps({k: v for k, v in doc.items() if k != "OBJECTS"}))
    assert p.solver == vtp._abi.SOLVER_IDS[solver]
    assert (p.time_step, p.particle_radius) == (np.float32(req["timeStep"]), np.float32(req["particleRadius"]))
    assert list(p.sim_space_size) == [np.float32(v) for v in req["simSpaceSize"]]
    # Optional.enable is false in every example: the solvers keep (0, -9.8, 0)
    assert opt["enable"] is False
    assert list(p.gravity) == [0.0, np.float32(-9.8), 0.0]
    if solver == "PBF":
        assert p.iteration_num == req["iterationNum"] and p.xsph_k == np.float32(req["XSPH_k"])
        assert p.b_rest_density == np.float32(req["bPartRestDensity"])
    if solver != "PBF":
        assert p.incomp_threshold == np.float32(req["incompThreshold"])
        assert p.surface_tension_coefficient == np.float32(req["surfaceTensionCoefficient"]) == np.float32(0.001)
    if solver in ("IMM", "IMMCT"):
        assert p.phase_num == len(req["phaseRestDensity"])
        assert list(p.phase_rest_density)[:p.phase_num] == [np.float32(v) for v in req["phaseRestDensity"]]
    if solver == "IMMCT":
        # the reference IMM-CT: two phases, polymer phase 1, solventViscosity for both phases
        assert p.phase_num == 2 and p.polymer_phase == 1 and "Extensions" not in doc[solver]
        assert list(p.phase_rest_viscosity)[:2] == [np.float32(req["solventViscosity"])] * 2
        assert p.solution_vis_base == np.float32(req["solutionBasicViscosity"])


@pytest.mark.skipif(not RUNNER.exists(), reason="host examples not built (make -C vt-physics_amd)")
@pytest.mark.parametrize("solver", sorted(SOLVERS))
def test_example_scene_objects_generate_identically(tmp_path, solver):
    doc = _load(f"Examples/{solver}/Scene-1/Configuration.json")
    sp = tmp_path / "scene.json"
    sp.write_text(json.dumps({"OBJECTS": doc["OBJECTS"]}))
    out = tmp_path / "gen"
    out.mkdir()
    r = subprocess.run([str(RUNNER), str(sp), "--generate-only", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    wl = scenes.from_scene_document(doc)
    counts = json.loads(r.stdout.strip().splitlines()[-1])["objects"]
    assert counts == [o.pos.shape[0] for o in wl.objects]
    for i, o in enumerate(wl.objects):
        cpp = np.fromfile(out / f"{i}.bin", dtype=np.float32).reshape(-1, 3)
        np.testing.assert_array_equal(cpp, o.pos)
    # cube [-1,1]^3 at spacing 0.1: 20^3 fluid particles in every example
    assert wl.objects[0].pos.shape[0] == 8000 and wl.objects[0].mat == vtp.MAT_FLUID
    assert wl.objects[1].mat == vtp.MAT_BOUNDARY
