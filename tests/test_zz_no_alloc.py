"""vtp_pbf_density_loss_noalloc (no allocation; scratch in a fixed device
array) against vtp_pbf_density_loss (allocates its scratch per call), through
the C ABI on cuda:0.

Two calls in a row on the same GPU, the second on a smaller context, so the
shared scratch holds the first call's values when the second starts.  The
gradient comes from the same kernel and the same per-particle dL/drho, so it
is bit-identical; the loss is an FP64 sum of per-wavefront atomics whose order
varies between launches, so it agrees to 1e-12 relative.  The outputs go to
device buffers allocated before the call with hipMalloc of the HIP runtime
libvtp.so itself runs on (the one loaded under the soname libamdhip64.so.7),
so the test does not depend on which runtime a process loaded first.
"""
import ctypes as C

import numpy as np
import pytest

import vtp
from vtp import scenes

pytestmark = pytest.mark.gpu


def _context(n_side, seed, ticks=3):
    wl = scenes.cfg4_adjoint(seed=seed, n_side=n_side)
    p = vtp.params_from_json("PBF", wl.config_json())
    ctx = vtp.Context(p)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, None, o.scene)
    ctx.initialize()
    ctx.step(ticks)
    return wl, ctx


def _hip():
    lib = C.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    lib.hipFree.argtypes = [C.c_void_p]
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


def _noalloc(ctx, n, with_grad=True):
    """Calls the entry point with device outputs; returns (status, loss, grad or None) on the host."""
    hip = _hip()
    d_loss, d_grad = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_loss), 8) == 0
    if with_grad:
        assert hip.hipMalloc(C.byref(d_grad), n * 12) == 0
    try:
        st = ctx.lib.vtp_pbf_density_loss_noalloc(ctx.h, d_loss, d_grad if with_grad else None)
        assert hip.hipDeviceSynchronize() == 0
        loss = np.zeros(1, np.float64)
        grad = np.zeros((n, 3), np.float32) if with_grad else None
        if st == vtp._abi.VTP_OK:
            assert hip.hipMemcpy(loss.ctypes.data, d_loss, 8, 2) == 0          # device to host
            if with_grad:
                assert hip.hipMemcpy(grad.ctypes.data, d_grad, n * 12, 2) == 0
        return st, float(loss[0]), grad
    finally:
        hip.hipFree(d_loss)
        if with_grad:
            hip.hipFree(d_grad)


def test_noalloc_density_loss_matches_the_allocating_entry_point():
    for n_side, seed in (((20, 16, 16), 0), ((10, 8, 8), 1)):   # the second input is smaller
        wl, ctx = _context(n_side, seed)
        print(f"density loss, noalloc vs allocating: {wl.n} particles", flush=True)
        L, g = ctx.density_loss()
        st, d_loss, d_grad = _noalloc(ctx, wl.n)
        assert st == vtp._abi.VTP_OK, ctx.lib.vtp_last_error(ctx.h)
        assert L > 0
        assert abs(d_loss - L) <= 1e-12 * abs(L)
        np.testing.assert_array_equal(d_grad, g)
        st, d_loss2, _ = _noalloc(ctx, wl.n, with_grad=False)
        assert st == vtp._abi.VTP_OK
        assert abs(d_loss2 - L) <= 1e-12 * abs(L)
        ctx.close()


def test_noalloc_density_loss_rejects_more_particles_than_its_scratch():
    n = (1 << 22) + 1
    rng = np.random.default_rng(0)
    pos = rng.uniform(0.0, 4.0, size=(n, 3)).astype(np.float32)
    cfg = scenes.pbf_config(lb=(-0.1, -0.1, -0.1), size=(4.2, 4.2, 4.2))
    ctx = vtp.Context(vtp.params_from_json("PBF", __import__("json").dumps(cfg)))
    ctx.add_particles(pos, vtp.MAT_FLUID)
    ctx.initialize()
    print(f"density loss noalloc, {n} particles: refused", flush=True)
    st, _, _ = _noalloc(ctx, n, with_grad=False)
    assert st == vtp._abi.VTP_EINVAL
    assert b"exceed its scratch" in ctx.lib.vtp_last_error(ctx.h)
    L, _ = ctx.density_loss(with_grad=False)   # the allocating entry point takes it
    assert np.isfinite(L)
    ctx.close()
