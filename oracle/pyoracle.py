"""ctypes binding of the FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_D = C.POINTER(C.c_double)
_U = C.POINTER(C.c_uint32)
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise OSError(f"{LIB} not built; run `make -C oracle`")
        lib = C.CDLL(str(LIB))
        lib.orc_create.restype = C.c_void_p
        lib.orc_create.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        for name in ("orc_initialize",):
            getattr(lib, name).argtypes = [C.c_void_p]
        lib.orc_step.argtypes = [C.c_void_p, C.c_uint32]
        lib.orc_read.argtypes = [C.c_void_p, C.c_int32, _D]
        lib.orc_write.argtypes = [C.c_void_p, C.c_int32, _D]
        lib.orc_neighbors.argtypes = [C.c_void_p, _U, _U, _U, _U, _U]
        lib.orc_search.argtypes = [C.c_void_p]
        lib.orc_set_stop.argtypes = [C.c_void_p, C.c_int]
        lib.orc_set_stop.restype = None
        lib.orc_stats.argtypes = [C.c_void_p, _U, _U, _U, _D, _D, _U]
        lib.orc_destroy.argtypes = [C.c_void_p]
        lib.orc_destroy.restype = None
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_set_threads.restype = None
        lib.orc_get_threads.restype = C.c_int
        lib.orc_pbf_rollout_fwd.argtypes = [C.c_void_p, C.c_uint32]
        lib.orc_pbf_rollout_bwd.argtypes = [C.c_void_p, _D, _D, _D, _D, _D]
        lib.orc_pbf_density_loss.argtypes = [C.c_void_p, _D, _D]
        lib.orc_pbf_set_xsph.argtypes = [C.c_void_p, C.c_double]
        lib.orc_kernel.argtypes = [C.c_double, _D, _D, _D, _D]
        lib.orc_kernel.restype = None
        lib.orc_cell_coord.argtypes = [C.c_float, C.c_float, C.c_float]
        lib.orc_neighbor_test.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_float]
        lib.orc_ct_update.argtypes = [_D, _D] + [C.c_double] * 8 + [_D, _D, _D]
        lib.orc_ct_update.restype = None
        _lib = lib
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _up(a):
    return None if a is None else a.ctypes.data_as(_U)


FIELD_COMPS = {0: 3, 1: 3, 14: 3, 15: 3, 16: 3, 20: 3, 22: 9, 23: 9, 24: 9}
# probes of internal IMM-CT arrays (oracle.h)
PROBE_BETA, PROBE_KAPPA_CT, PROBE_BLOCKING = 100, 101, 102


class Oracle:
    """FP64 reference simulation in attach order."""

    def __init__(self, params, pos, vel=None, vol_frac=None, scene=None, objects=None):
        self.lib = load()
        self.params = params
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        self.n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(-1, 3)
        self.P = params.phase_num if params.solver in (2, 3) else 0
        vf = None if vol_frac is None else np.ascontiguousarray(vol_frac, np.float32)
        sc = None if scene is None else np.ascontiguousarray(scene, np.uint32)
        st, en, mt = objects
        st = np.ascontiguousarray(st, np.uint32)
        en = np.ascontiguousarray(en, np.uint32)
        mt = np.ascontiguousarray(mt, np.int32)
        self._keep = (pos, vel, vf, sc, st, en, mt)
        self.h = self.lib.orc_create(C.byref(params), self.n, pos.ctypes.data, None if vel is None else vel.ctypes.data,
                                     None if vf is None else vf.ctypes.data, None if sc is None else sc.ctypes.data,
                                     len(st), st.ctypes.data, en.ctypes.data, mt.ctypes.data)
        if not self.h:
            raise RuntimeError("orc_create failed")

    @classmethod
    def from_workload(cls, params, wl):
        pos, vel, frac, scene, objs = wl.concat()
        return cls(params, pos, vel, frac, scene, objs)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def initialize(self):
        self.lib.orc_initialize(self.h)

    def step(self, n=1):
        return self.lib.orc_step(self.h, n)

    def read(self, fid: int, comps: int | None = None):
        if comps is None:
            comps = FIELD_COMPS.get(fid, 1)
            if fid == 18:
                comps = self.P
            if fid in (19, 25):
                comps = 3 * self.P
        out = np.empty(self.n * comps, np.float64)
        rc = self.lib.orc_read(self.h, fid, _dp(out))
        if rc != 0:
            raise RuntimeError(f"orc_read({fid}) failed: {rc}")
        return out.reshape(self.n, comps) if comps > 1 else out

    def write(self, fid: int, arr):
        a = np.ascontiguousarray(arr, np.float64).reshape(-1)
        rc = self.lib.orc_write(self.h, fid, _dp(a))
        if rc != 0:
            raise RuntimeError("orc_write failed")

    def neighbors(self, cells):
        M = self.params.max_neighbor_num
        lists = np.empty(self.n * M, np.uint32)
        num = np.empty(self.n, np.uint32)
        srt = np.empty(self.n, np.uint32)
        cs = np.empty(cells, np.uint32)
        ce = np.empty(cells, np.uint32)
        self.lib.orc_neighbors(self.h, _up(lists), _up(num), _up(srt), _up(cs), _up(ce))
        return {"neighbors": lists.reshape(self.n, M), "neighbor_num": num, "sorted": srt,
                "cell_start": cs, "cell_end": ce}

    def set_stop(self, stage: int):
        """Test probe: 1 = a DFSPH tick returns right after its divergence solve."""
        self.lib.orc_set_stop(self.h, stage)

    def search(self):
        """Neighbour search of the current positions alone (no step)."""
        if self.lib.orc_search(self.h) != 0:
            raise RuntimeError("orc_search failed")

    def stats(self):
        a, b, c, e = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        d1, d2 = C.c_double(), C.c_double()
        self.lib.orc_stats(self.h, C.byref(a), C.byref(b), C.byref(c), C.byref(d1), C.byref(d2), C.byref(e))
        return {"div_iterations": a.value, "incomp_iterations": b.value, "phase_iterations": c.value,
                "div_residual": d1.value, "incomp_residual": d2.value, "crash_flags": e.value}

    def rollout_fwd(self, nsteps):
        rc = self.lib.orc_pbf_rollout_fwd(self.h, nsteps)
        if rc != 0:
            raise RuntimeError(f"orc_pbf_rollout_fwd failed: {rc}")

    def rollout_bwd(self, grad_x=None, grad_v=None):
        gx = None if grad_x is None else np.ascontiguousarray(grad_x, np.float64)
        gv = None if grad_v is None else np.ascontiguousarray(grad_v, np.float64)
        gv0 = np.zeros((self.n, 3))
        gx0 = np.zeros((self.n, 3))
        gk = np.zeros(1)
        rc = self.lib.orc_pbf_rollout_bwd(self.h, _dp(gx), _dp(gv), _dp(gv0), _dp(gx0), _dp(gk))
        if rc != 0:
            raise RuntimeError(f"orc_pbf_rollout_bwd failed: {rc}")
        return gv0, gx0, float(gk[0])

    def density_loss(self):
        loss = np.zeros(1)
        g = np.zeros((self.n, 3))
        rc = self.lib.orc_pbf_density_loss(self.h, _dp(loss), _dp(g))
        if rc != 0:
            raise RuntimeError("orc_pbf_density_loss failed")
        return float(loss[0]), g

    def set_xsph(self, k):
        self.lib.orc_pbf_set_xsph(self.h, float(k))


def kernel(h: float, r):
    """(W, gradW, Akinci C) of the oracle's kernel functions at offset r."""
    lib = load()
    rr = np.ascontiguousarray(r, np.float64)
    w = C.c_double()
    g = np.zeros(3)
    c = C.c_double()
    lib.orc_kernel(h, _dp(rr), C.byref(w), _dp(g), C.byref(c))
    return w.value, g, c.value


def ct_update(Q, L, dt, relax, beta, kappa, vis_base, vis_max, a1, th):
    """One conformation update of the oracle (update_ct_cuda): (Q_new, mu, stress)."""
    lib = load()
    q = np.ascontiguousarray(Q, np.float64).reshape(9)
    l = np.ascontiguousarray(L, np.float64).reshape(9)
    qo = np.zeros(9)
    st = np.zeros(9)
    mu = C.c_double()
    lib.orc_ct_update(_dp(q), _dp(l), dt, relax, beta, kappa, vis_base, vis_max, a1, th, _dp(qo), C.byref(mu),
                      _dp(st))
    return qo.reshape(3, 3), mu.value, st.reshape(3, 3)


def set_threads(n: int):
    load().orc_set_threads(n)


def get_threads() -> int:
    return load().orc_get_threads()
