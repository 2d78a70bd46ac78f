"""ctypes binding of the C ABI declared in include/vtp.h.

The shared library libvtp.so is built in-tree (vt-physics_amd/libvtp.so).  It
is the only compute path: nothing here falls back to a CPU implementation.
If the library is missing the import fails loudly.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "libvtp.so"

VTP_OK, VTP_EINVAL, VTP_ENOMEM, VTP_EHIP, VTP_ECRASH, VTP_ESTATE, VTP_ENODEV, VTP_ENBRCAP = range(8)
STATUS_NAMES = {0: "VTP_OK", 1: "VTP_EINVAL", 2: "VTP_ENOMEM", 3: "VTP_EHIP", 4: "VTP_ECRASH",
                5: "VTP_ESTATE", 6: "VTP_ENODEV", 7: "VTP_ENBRCAP"}
PBF, DFSPH, IMM, IMMCT = 0, 1, 2, 3
SOLVER_IDS = {"PBF": PBF, "DFSPH": DFSPH, "IMM": IMM, "IMMCT": IMMCT}
MAT_FLUID, MAT_BOUNDARY = 0, 1
MAX_PHASES = 8

CRASH_DIV_ITER_CAP = 1 << 0
CRASH_INCOMP_ITER_CAP = 1 << 1
CRASH_PHASE_ITER_CAP = 1 << 2
CRASH_NAN = 1 << 3
CRASH_OUT_OF_GRID = 1 << 4

FIELDS = {
    "pos": (0, 3), "vel": (1, 3), "mat": (2, 1), "mass": (3, 1), "volume": (4, 1), "density": (5, 1),
    "lambda": (6, 1), "error": (7, 1), "rest_density": (8, 1), "compression_ratio": (9, 1),
    "df_alpha": (10, 1), "delta_cr": (11, 1), "kappa_div": (12, 1), "kappa_incomp": (13, 1),
    "vel_adv": (14, 3), "acc": (15, 3), "surface_normal": (16, 3), "vis": (17, 1), "vol_frac": (18, "P"),
    "vel_phase": (19, "P3"), "color": (20, 3), "cd": (21, 1), "Q": (22, 9), "vel_grad": (23, 9),
    "vise_stress": (24, 9), "vel_drift": (25, "P3"), "scene": (26, 1),
}
INT_FIELDS = {"mat": np.int32, "scene": np.uint32}


class VtpParams(C.Structure):
    _fields_ = [
        ("solver", C.c_int32),
        ("animation_time", C.c_float),
        ("time_step", C.c_float),
        ("particle_radius", C.c_float),
        ("sim_space_lb", C.c_float * 3),
        ("sim_space_size", C.c_float * 3),
        ("max_neighbor_num", C.c_uint32),
        ("gravity", C.c_float * 3),
        ("iteration_num", C.c_uint32),
        ("xsph_k", C.c_float),
        ("f_rest_density", C.c_float),
        ("b_rest_density", C.c_float),
        ("f_rest_viscosity", C.c_float),
        ("div_free_threshold", C.c_float),
        ("incomp_threshold", C.c_float),
        ("surface_tension_coefficient", C.c_float),
        ("phase_num", C.c_uint32),
        ("phase_rest_density", C.c_float * MAX_PHASES),
        ("phase_rest_viscosity", C.c_float * MAX_PHASES),
        ("phase_rest_color", (C.c_float * 3) * MAX_PHASES),
        ("cf", C.c_float),
        ("cd", C.c_float),
        ("solution_vis_base", C.c_float),
        ("solution_vis_max", C.c_float),
        ("relaxation_time", C.c_float),
        ("thinning_exp0", C.c_float),
        ("rheology_threshold", C.c_float),
        ("polymer_phase", C.c_int32),
        ("cfl_enable", C.c_int32),
        ("cfl_number", C.c_float),
        ("dt_max", C.c_float),
        ("max_solver_iterations", C.c_uint32),
        ("mass_weighted", C.c_int32),
    ]


class VtpStepStats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("sim_time", C.c_float),
        ("dt", C.c_float),
        ("div_iterations", C.c_uint32),
        ("incomp_iterations", C.c_uint32),
        ("phase_iterations", C.c_uint32),
        ("pad_", C.c_uint32),
        ("div_residual", C.c_double),
        ("incomp_residual", C.c_double),
        ("total_div_iterations", C.c_uint64),
        ("total_incomp_iterations", C.c_uint64),
        ("dead_particles", C.c_uint32),
        ("crash_flags", C.c_uint32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad_"}


class VtpStateHeader(C.Structure):
    """vtp_state_header: the fixed header of a state image."""
    _fields_ = [
        ("magic", C.c_char * 8),
        ("version", C.c_uint32),
        ("solver", C.c_int32),
        ("n", C.c_uint32),
        ("phases", C.c_uint32),
        ("objects", C.c_uint32),
        ("crash_flags", C.c_uint32),
        ("layout_hash", C.c_uint64),
        ("steps", C.c_uint64),
        ("sim_time", C.c_double),
        ("total_div_iterations", C.c_uint64),
        ("total_incomp_iterations", C.c_uint64),
        ("iter_hint_div", C.c_uint32),
        ("iter_hint_incomp", C.c_uint32),
        ("nbr_at_cap", C.c_uint32),
        ("reserved", C.c_uint32),
        ("payload_bytes", C.c_uint64),
    ]


def state_sections(header: VtpStateHeader) -> list[tuple[str, np.dtype, tuple]]:
    """(name, dtype, shape) of each array that follows the header, in order."""
    n, P = header.n, header.phases
    out = [("orig", np.uint32, (n,)), ("x", np.float32, (n, 4)), ("v", np.float32, (n, 4))]
    if header.solver in (IMM, IMMCT):
        out += [("vol_frac", np.float32, (P, n)), ("vel_phase", np.float32, (P, n, 4)),
                ("rest_density", np.float32, (n,)), ("color", np.float32, (n, 4))]
    if header.solver == IMMCT:
        out.append(("Q_minus_I", np.float32, (9, n)))
    return out


def parse_state(data: bytes) -> tuple[VtpStateHeader, dict]:
    """Split a state image into its header and named arrays (views, no copy)."""
    hs = C.sizeof(VtpStateHeader)
    if len(data) < hs:
        raise ValueError("state image shorter than its header")
    h = VtpStateHeader.from_buffer_copy(data[:hs])
    if h.magic != b"VTPSTATE":
        raise ValueError("not a state image")
    arrays, off = {}, hs
    for name, dt, shape in state_sections(h):
        cnt = int(np.prod(shape))
        arrays[name] = np.frombuffer(data, dtype=dt, count=cnt, offset=off).reshape(shape)
        off += cnt * np.dtype(dt).itemsize
    if off != len(data) or off - hs != h.payload_bytes:
        raise ValueError("state image size does not match its header")
    return h, arrays


# Every symbol include/vtp.h declares, with (restype, argtypes).
_P = C.c_void_p
_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
SIGNATURES = {
    "vtp_params_from_json": (C.c_int, [C.c_int32, C.c_char_p, C.POINTER(VtpParams), C.c_char_p, C.c_size_t]),
    "vtp_params_default": (C.c_int, [C.c_int32, C.POINTER(VtpParams)]),
    "vtp_config_template": (C.c_char_p, [C.c_int32]),
    "vtp_object_config_template": (C.c_char_p, [C.c_int32]),
    "vtp_create": (C.c_int, [C.POINTER(_P), C.c_int32, C.c_char_p]),
    "vtp_create_with_params": (C.c_int, [C.POINTER(_P), C.POINTER(VtpParams)]),
    "vtp_add_particles": (C.c_int, [_P, C.c_uint32, C.c_int32, _F, _F, _F, C.c_uint32]),
    "vtp_initialize": (C.c_int, [_P]),
    "vtp_step": (C.c_int, [_P, C.c_uint32]),
    "vtp_run": (C.c_int, [_P]),
    "vtp_destroy": (None, [_P]),
    "vtp_last_error": (C.c_char_p, [_P]),
    "vtp_particle_count": (C.c_uint32, [_P]),
    "vtp_read": (C.c_int, [_P, C.c_int32, C.c_void_p, C.c_size_t]),
    "vtp_write": (C.c_int, [_P, C.c_int32, C.c_void_p, C.c_size_t]),
    "vtp_read_device": (C.c_int, [_P, C.c_int32, C.c_void_p, C.c_size_t]),
    "vtp_write_device": (C.c_int, [_P, C.c_int32, C.c_void_p, C.c_size_t]),
    "vtp_stats": (C.c_int, [_P, C.POINTER(VtpStepStats)]),
    "vtp_debug_neighbors": (C.c_int, [_P, _U, _U, _U, _U, _U]),
    "vtp_grid_info": (C.c_int, [_P, _U, _U]),
    "vtp_state_size": (C.c_int, [_P, C.POINTER(C.c_size_t)]),
    "vtp_save_state": (C.c_int, [_P, C.c_void_p, C.c_size_t]),
    "vtp_load_state": (C.c_int, [_P, C.c_void_p, C.c_size_t]),
    "vtp_set_option": (C.c_int, [_P, C.c_char_p, C.c_double]),
    "vtp_get_option": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double)]),
    "vtp_stage_times": (C.c_int32, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_double),
                                    C.POINTER(C.c_uint64), C.c_int32]),
    "vtp_reset_stage_times": (C.c_int, [_P]),
    "vtp_pbf_rollout_fwd": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "vtp_pbf_rollout_bwd": (C.c_int, [_P, _F, _F, _F, _F, _F]),
    "vtp_pbf_rollout_bwd_device": (C.c_int, [_P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "vtp_pbf_density_loss_device": (C.c_int, [_P, C.c_void_p, C.c_void_p]),
    "vtp_pbf_density_loss": (C.c_int, [_P, C.POINTER(C.c_double), _F]),
    "vtp_pbf_density_loss_noalloc": (C.c_int, [_P, C.c_void_p, C.c_void_p]),
    "vtp_pbf_set_xsph": (C.c_int, [_P, C.c_float]),
    "vtp_version": (C.c_char_p, []),
}

_lib = None


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libvtp.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise OSError(f"libvtp.so not found at {p}; build it with `make -C vt-physics_amd` "
                      "or __graft_entry__.build()")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class VtpError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_F)


def uptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_U)


def params_from_json(solver: int | str, text: str) -> VtpParams:
    lib = load_library()
    sid = SOLVER_IDS[solver] if isinstance(solver, str) else solver
    p = VtpParams()
    err = C.create_string_buffer(512)
    st = lib.vtp_params_from_json(sid, text.encode(), C.byref(p), err, 512)
    if st != VTP_OK:
        raise VtpError(st, err.value.decode())
    return p


def params_default(solver: int | str) -> VtpParams:
    lib = load_library()
    sid = SOLVER_IDS[solver] if isinstance(solver, str) else solver
    p = VtpParams()
    st = lib.vtp_params_default(sid, C.byref(p))
    if st != VTP_OK:
        raise VtpError(st, "vtp_params_default failed")
    return p


def config_template(solver: int | str) -> dict:
    """The solver's config template (Solver::getSolverConfigTemplate)."""
    sid = SOLVER_IDS[solver] if isinstance(solver, str) else solver
    txt = load_library().vtp_config_template(sid)
    if txt is None:
        raise VtpError(VTP_EINVAL, f"unknown solver {solver!r}")
    return json.loads(txt.decode())


def object_config_template(solver: int | str) -> dict:
    """Solver::getSolverObjectComponentConfigTemplate."""
    sid = SOLVER_IDS[solver] if isinstance(solver, str) else solver
    txt = load_library().vtp_object_config_template(sid)
    if txt is None:
        raise VtpError(VTP_EINVAL, f"unknown solver {solver!r}")
    return json.loads(txt.decode())


class Context:
    """One solver instance behind the C ABI (owns the device state)."""

    def __init__(self, params: VtpParams):
        self.lib = load_library()
        self.params = params
        self.P = params.phase_num if params.solver in (IMM, IMMCT) else 0
        h = C.c_void_p()
        st = self.lib.vtp_create_with_params(C.byref(h), C.byref(params))
        if st != VTP_OK:
            raise VtpError(st, self.lib.vtp_last_error(None).decode())
        self.h = h
        self.n = 0

    @classmethod
    def from_json(cls, solver: int | str, text: str) -> "Context":
        return cls(params_from_json(solver, text))

    def _check(self, st: int):
        if st != VTP_OK:
            raise VtpError(st, self.lib.vtp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.vtp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_particles(self, pos, mat=MAT_FLUID, vel=None, vol_frac=None, scene=0):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        f = None if vol_frac is None else np.ascontiguousarray(vol_frac, dtype=np.float32).reshape(n, -1)
        self._check(self.lib.vtp_add_particles(self.h, scene, mat, fptr(pos), fptr(v), fptr(f), n))
        self.n += n

    def initialize(self):
        self._check(self.lib.vtp_initialize(self.h))

    def step(self, n: int = 1):
        self._check(self.lib.vtp_step(self.h, n))

    def run(self):
        self._check(self.lib.vtp_run(self.h))

    def set_option(self, key: str, value: float):
        self._check(self.lib.vtp_set_option(self.h, key.encode(), float(value)))

    def get_option(self, key: str) -> float:
        v = C.c_double()
        self._check(self.lib.vtp_get_option(self.h, key.encode(), C.byref(v)))
        return v.value

    TUNED_PASSES = {0: ("pbf_lambda", "pbf_dx", "pbf_xsph"),
                    1: ("df_precompute", "df_delta_kappa", "df_update_vadv", "df_vis_normal"),
                    2: ("df_delta_kappa", "df_update_vadv", "imm_precompute", "imm_vis_normal",
                        "imm_tension_phase_vel", "imm_fluxes"),
                    3: ("df_delta_kappa", "df_update_vadv", "imm_precompute", "imm_vis_normal",
                        "imm_tension_phase_vel", "imm_fluxes", "immct_conformation", "immct_stress_acc")}

    def launch_shapes(self, solver_id: int) -> dict:
        """Workgroup size of every neighbour pass of the solver, and the PBF lanes."""
        d = {p: int(self.get_option("block." + p)) for p in self.TUNED_PASSES[solver_id]}
        d["lanes"] = int(self.get_option("lanes"))
        return d

    def _shape(self, field: str):
        fid, comps = FIELDS[field]
        n = self.lib.vtp_particle_count(self.h)
        if comps == "P":
            return fid, (n, self.P)
        if comps == "P3":
            return fid, (n, self.P, 3)
        return fid, (n,) if comps == 1 else (n, comps)

    def read(self, field: str) -> np.ndarray:
        fid, shape = self._shape(field)
        out = np.empty(shape, dtype=INT_FIELDS.get(field, np.float32))
        self._check(self.lib.vtp_read(self.h, fid, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def write(self, field: str, arr):
        fid, shape = self._shape(field)
        a = np.ascontiguousarray(arr, dtype=np.float32).reshape(shape)
        self._check(self.lib.vtp_write(self.h, fid, a.ctypes.data_as(C.c_void_p), a.nbytes))

    # device-pointer variants: `ptr` is the address of a float[n][3] array in
    # attach order on the context's GPU (e.g. torch's data_ptr()); the caller
    # has finished writing it, and the call returns after the context's stream
    def read_device(self, field: str, ptr: int):
        fid, shape = self._shape(field)
        self._check(self.lib.vtp_read_device(self.h, fid, ptr, 12 * shape[0]))

    def write_device(self, field: str, ptr: int):
        fid, shape = self._shape(field)
        self._check(self.lib.vtp_write_device(self.h, fid, ptr, 12 * shape[0]))

    def stats(self) -> dict:
        s = VtpStepStats()
        self._check(self.lib.vtp_stats(self.h, C.byref(s)))
        return s.as_dict()

    def grid_info(self):
        g = (C.c_uint32 * 3)()
        cells = C.c_uint32()
        self._check(self.lib.vtp_grid_info(self.h, g, C.byref(cells)))
        return tuple(g), cells.value

    def debug_neighbors(self, scenes: int = 1):
        n = self.lib.vtp_particle_count(self.h)
        M = self.params.max_neighbor_num
        _, cells = self.grid_info()
        lists = np.empty(n * M, np.uint32)
        num = np.empty(n, np.uint32)
        srt = np.empty(n, np.uint32)
        cs = np.empty(cells * scenes, np.uint32)
        ce = np.empty(cells * scenes, np.uint32)
        self._check(self.lib.vtp_debug_neighbors(self.h, uptr(lists), uptr(num), uptr(srt), uptr(cs), uptr(ce)))
        return {"neighbors": lists.reshape(n, M), "neighbor_num": num, "sorted": srt,
                "cell_start": cs, "cell_end": ce}

    def save_state(self, path: os.PathLike | str | None = None) -> bytes:
        """State image of the context (vtp_save_state); also written to `path` if given."""
        size = C.c_size_t()
        self._check(self.lib.vtp_state_size(self.h, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        self._check(self.lib.vtp_save_state(self.h, buf, size.value))
        data = buf.raw
        if path is not None:
            Path(path).write_bytes(data)
        return data

    def load_state(self, data: bytes | os.PathLike | str):
        """Resume from a state image (bytes, or the path of a file save_state wrote)."""
        if not isinstance(data, (bytes, bytearray)):
            data = Path(data).read_bytes()
        buf = C.create_string_buffer(bytes(data), len(data))
        self._check(self.lib.vtp_load_state(self.h, buf, len(data)))

    def stage_times(self):
        names = C.create_string_buffer(8192)
        ms = (C.c_double * 64)()
        cnt = (C.c_uint64 * 64)()
        k = self.lib.vtp_stage_times(self.h, names, 8192, ms, cnt, 64)
        if k < 0:
            raise VtpError(VTP_EHIP, self.lib.vtp_last_error(self.h).decode())
        keys = [s for s in names.value.decode().split("\n") if s]
        return {keys[i]: (ms[i], cnt[i]) for i in range(min(k, 64))}

    def reset_stage_times(self):
        self._check(self.lib.vtp_reset_stage_times(self.h))

    # differentiable PBF rollout
    def rollout_fwd(self, nsteps: int, ckpt_every: int):
        self._check(self.lib.vtp_pbf_rollout_fwd(self.h, nsteps, ckpt_every))

    def rollout_bwd(self, grad_x=None, grad_v=None):
        n = self.lib.vtp_particle_count(self.h)
        gx = None if grad_x is None else np.ascontiguousarray(grad_x, np.float32).reshape(n, 3)
        gv = None if grad_v is None else np.ascontiguousarray(grad_v, np.float32).reshape(n, 3)
        gv0 = np.zeros((n, 3), np.float32)
        gx0 = np.zeros((n, 3), np.float32)
        gk = np.zeros(1, np.float32)
        self._check(self.lib.vtp_pbf_rollout_bwd(self.h, fptr(gx), fptr(gv), fptr(gv0), fptr(gx0), fptr(gk)))
        return gv0, gx0, float(gk[0])

    def density_loss(self, with_grad: bool = True):
        n = self.lib.vtp_particle_count(self.h)
        loss = C.c_double()
        g = np.zeros((n, 3), np.float32) if with_grad else None
        self._check(self.lib.vtp_pbf_density_loss(self.h, C.byref(loss), fptr(g)))
        return loss.value, g

    def rollout_bwd_device(self, gx_ptr, gv_ptr, gv0_ptr, gx0_ptr, gk_ptr):
        """vtp_pbf_rollout_bwd on device pointers (0 or None for an absent array)."""
        self._check(self.lib.vtp_pbf_rollout_bwd_device(self.h, gx_ptr or None, gv_ptr or None, gv0_ptr or None,
                                                        gx0_ptr or None, gk_ptr or None))

    def density_loss_device(self, loss_ptr, g_ptr=None):
        """vtp_pbf_density_loss with the loss (one double) and dL/dx in device memory."""
        self._check(self.lib.vtp_pbf_density_loss_device(self.h, loss_ptr, g_ptr or None))

    def set_xsph(self, k: float):
        self._check(self.lib.vtp_pbf_set_xsph(self.h, float(k)))
