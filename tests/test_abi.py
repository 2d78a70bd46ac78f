"""C ABI surface and configuration parsing (host only, no GPU needed)."""
import ctypes as C
import json
import re
from pathlib import Path

import numpy as np
import pytest

import vtp
from vtp import _abi

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "vtp.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vtp_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(lib):
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), f"libvtp.so does not export {name}"
    assert set(names) == set(_abi.SIGNATURES), "ctypes signature table out of sync with include/vtp.h"


def test_version_string(lib):
    assert b"gfx950" in lib.vtp_version()


def test_create_without_gpu_fails_cleanly(lib):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    p = vtp.params_default("PBF")
    h = C.c_void_p()
    st = lib.vtp_create_with_params(C.byref(h), C.byref(p))
    assert st == _abi.VTP_ENODEV
    assert h.value is None
    assert b"device" in lib.vtp_last_error(None)


def test_null_context_calls_are_rejected(lib):
    assert lib.vtp_step(None, 1) == _abi.VTP_EINVAL
    assert lib.vtp_initialize(None) == _abi.VTP_EINVAL
    assert lib.vtp_particle_count(None) == 0
    lib.vtp_destroy(None)


@pytest.mark.parametrize("solver", ["PBF", "DFSPH", "IMM", "IMMCT"])
def test_template_defaults(solver):
    p = vtp.params_default(solver)
    assert p.solver == _abi.SOLVER_IDS[solver]
    assert p.max_neighbor_num == 60
    # gravity is (0,-9.8,0) unless Optional.enable (PBFSolver.cpp:53-58); templates have enable=false
    assert list(p.gravity) == [0.0, np.float32(-9.8), 0.0]
    assert p.max_solver_iterations == 100


def test_pbf_template_values():
    p = vtp.params_default("PBF")
    assert np.isclose(p.time_step, 0.01) and np.isclose(p.particle_radius, 0.075)
    assert p.iteration_num == 10 and np.isclose(p.xsph_k, 0.02)
    assert list(p.sim_space_lb) == [-1, -1, -1] and list(p.sim_space_size) == [2, 2, 2]


def test_immct_solvent_viscosity_fills_every_phase():
    p = vtp.params_default("IMMCT")
    assert p.phase_num == 2
    assert np.allclose(list(p.phase_rest_viscosity)[:2], [0.01, 0.01])
    assert np.isclose(p.solution_vis_base, 12) and np.isclose(p.relaxation_time, 0.001)
    assert p.polymer_phase == 1


def test_gravity_override_when_enabled():
    cfg = vtp.scenes.pbf_config()
    cfg["PBF"]["Optional"] = {"enable": True, "gravity": [1.0, 2.0, 3.0]}
    p = vtp.params_from_json("PBF", json.dumps(cfg))
    assert list(p.gravity) == [1.0, 2.0, 3.0]


@pytest.mark.parametrize("key", ["timeStep", "iterationNum", "XSPH_k", "bPartRestDensity"])
def test_missing_required_key_is_named(key):
    cfg = vtp.scenes.pbf_config()
    del cfg["PBF"]["Required"][key]
    with pytest.raises(vtp.VtpError, match=f"missing Required key: {key}"):
        vtp.params_from_json("PBF", json.dumps(cfg))


def test_missing_optional_key_is_an_error():
    cfg = vtp.scenes.pbf_config()
    del cfg["PBF"]["Optional"]["gravity"]
    with pytest.raises(vtp.VtpError, match="Optional key: gravity"):
        vtp.params_from_json("PBF", json.dumps(cfg))


def test_missing_domain_and_bad_json():
    with pytest.raises(vtp.VtpError, match="main domain: DFSPH"):
        vtp.params_from_json("DFSPH", json.dumps(vtp.scenes.pbf_config()))
    with pytest.raises(vtp.VtpError, match="parse error"):
        vtp.params_from_json("PBF", "{\"PBF\": [1, 2,}")


def test_phase_size_mismatch_rejected():
    cfg = vtp.scenes.immct_config(three_phase=False)
    cfg["IMMCT"]["Required"]["phaseRestColor"] = [[1, 2, 3]]
    with pytest.raises(vtp.VtpError, match="size not match"):
        vtp.params_from_json("IMMCT", json.dumps(cfg))


def test_three_phase_extension_keys():
    cfg = vtp.scenes.immct_config(three_phase=True)
    p = vtp.params_from_json("IMMCT", json.dumps(cfg))
    assert p.phase_num == 3 and p.polymer_phase == 1
    assert np.allclose(list(p.phase_rest_viscosity)[:3], [1e-3, 1e-3, 1e-2])
    assert np.allclose(list(p.phase_rest_density)[:3], [1000, 800, 1200])


def test_cfl_extension():
    cfg = vtp.scenes.dfsph_config(cfl=True)
    p = vtp.params_from_json("DFSPH", json.dumps(cfg))
    assert p.cfl_enable == 1 and np.isclose(p.cfl_number, 0.4) and np.isclose(p.dt_max, 0.002)


@pytest.mark.parametrize("solver", ["PBF", "IMM", "IMMCT"])
def test_cfl_extension_is_rejected_outside_dfsph(lib, solver):
    """Only the DFSPH tick adapts dt and advances the simulated time by it; a
    CFL flag on another solver would leave sim_time at 0 and vtp_run looping."""
    cfg = vtp.config_template(solver)
    dom = next(k for k in cfg if k != "EXPORT")
    cfg[dom]["Extensions"] = {"cflEnable": True}
    with pytest.raises(vtp.VtpError, match="cflEnable exists for DFSPH only"):
        vtp.params_from_json(solver, json.dumps(cfg))
    p = vtp.params_default(solver)
    p.cfl_enable = 1
    h = C.c_void_p()
    assert lib.vtp_create_with_params(C.byref(h), C.byref(p)) == _abi.VTP_EINVAL
    assert h.value is None
    assert b"DFSPH only" in lib.vtp_last_error(None)


def test_state_header_layout_matches_c(tmp_path):
    """ctypes VtpStateHeader has the C struct's size and field offsets."""
    import subprocess
    fields = [f for f, _ in _abi.VtpStateHeader._fields_]
    prog = tmp_path / "hdr.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "vtp.h"\nint main(void) {\n'
                    '  printf("%zu\\n", sizeof(vtp_state_header));\n'
                    + "".join(f'  printf("%zu\\n", offsetof(vtp_state_header, {f}));\n' for f in fields)
                    + "  return 0;\n}\n")
    exe = tmp_path / "hdr"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(prog), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(_abi.VtpStateHeader)] + [getattr(_abi.VtpStateHeader, f).offset for f in fields]
    assert got == want


def _image(solver, n, P):
    h = _abi.VtpStateHeader()
    h.magic = b"VTPSTATE"
    h.version, h.solver, h.n, h.phases, h.steps = 1, solver, n, P, 7
    rng = np.random.default_rng(0)
    parts = []
    for name, dt, shape in _abi.state_sections(h):
        a = rng.permutation(n).astype(np.uint32) if name == "orig" else rng.random(shape).astype(dt)
        parts.append((name, a))
    h.payload_bytes = sum(a.nbytes for _, a in parts)
    return bytes(h) + b"".join(a.tobytes() for _, a in parts), dict(parts)


@pytest.mark.parametrize("solver,P,names", [
    (_abi.PBF, 0, ["orig", "x", "v"]),
    (_abi.IMM, 3, ["orig", "x", "v", "vol_frac", "vel_phase", "rest_density", "color"]),
    (_abi.IMMCT, 2, ["orig", "x", "v", "vol_frac", "vel_phase", "rest_density", "color", "Q_minus_I"]),
])
def test_parse_state_sections(solver, P, names):
    data, parts = _image(solver, 37, P)
    h, arrays = vtp.parse_state(data)
    assert h.n == 37 and h.steps == 7 and list(arrays) == names
    for k in names:
        assert np.array_equal(arrays[k], parts[k])
    with pytest.raises(ValueError):
        vtp.parse_state(data[:-1])
    with pytest.raises(ValueError):
        vtp.parse_state(b"XXXXXXXX" + data[8:])


def test_state_calls_reject_null_context(lib):
    size = C.c_size_t()
    assert lib.vtp_state_size(None, C.byref(size)) == _abi.VTP_EINVAL
    assert lib.vtp_save_state(None, None, 0) == _abi.VTP_EINVAL
    assert lib.vtp_load_state(None, None, 0) == _abi.VTP_EINVAL
