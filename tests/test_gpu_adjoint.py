"""Hand-written PBF adjoint on the GPU (checkpointed rollout) against the
FP64 oracle adjoint and against central finite differences.

Tolerances: gradients with respect to the initial velocities and the XSPH
coefficient within 1e-3 relative (L2 norm for the velocity gradient) of the
FP64 adjoint, and directional derivatives along 16 seeded unit directions
within 1e-3 |grad| of central differences (eps 1e-3) of the FP64 forward
rollout; the XSPH derivative within 1e-3 relative.
"""
import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes

pytestmark = pytest.mark.gpu

STEPS, CKPT = 12, 5


def scene(seed=0):
    wl = scenes.cfg4_adjoint(seed=seed, n_side=(10, 8, 8))
    return wl


def setup(wl):
    p = vtp.params_from_json("PBF", wl.config_json())
    ctx = vtp.Context(p)
    for o in wl.objects:
        ctx.add_particles(o.pos, o.mat, o.vel, None, o.scene)
    ctx.initialize()
    return p, ctx


def loss_weights(wl, seed=1):
    rng = np.random.default_rng(seed)
    n, nf = wl.n, wl.n_fluid
    cx = rng.normal(size=(n, 3))
    cv = rng.normal(size=(n, 3))
    cx[nf:] = 0
    cv[nf:] = 0
    return cx, cv


def test_gpu_adjoint_matches_fp64_adjoint():
    wl = scene()
    p, ctx = setup(wl)
    cx, cv = loss_weights(wl)
    ctx.rollout_fwd(STEPS, CKPT)
    L, gL = ctx.density_loss()
    gv0, gx0, gk = ctx.rollout_bwd(cx + 100 * gL, cv)
    o = pyoracle.Oracle.from_workload(p, wl)
    o.initialize()
    o.rollout_fwd(STEPS)
    Lo, gLo = o.density_loss()
    assert abs(L - Lo) <= 1e-4 * abs(Lo)
    rv0, rx0, rk = o.rollout_bwd(cx + 100 * gLo, cv)
    nf = wl.n_fluid
    e = np.linalg.norm(gv0[:nf] - rv0[:nf]) / np.linalg.norm(rv0[:nf])
    assert e <= 1e-3, e
    assert abs(gk - rk) <= 1e-3 * abs(rk), (gk, rk)
    # the context is back at the rollout's final state
    np.testing.assert_allclose(ctx.read("pos")[:nf], o.read(0)[:nf], atol=1e-5)


def test_gpu_adjoint_directional_central_differences():
    wl = scene(seed=3)
    p, ctx = setup(wl)
    cx, cv = loss_weights(wl, seed=4)
    ctx.rollout_fwd(STEPS, CKPT)
    _, gL = ctx.density_loss()
    gv0, _, gk = ctx.rollout_bwd(cx + 100 * gL, cv)
    n, nf = wl.n, wl.n_fluid
    v_init = np.concatenate([o.vel if o.vel is not None else np.zeros_like(o.pos) for o in wl.objects])

    def loss(vel0=None, k=None):
        o = pyoracle.Oracle.from_workload(p, wl)
        o.initialize()
        if vel0 is not None:
            o.write(1, vel0)
        if k is not None:
            o.set_xsph(k)
        o.step(STEPS)
        Ld, _ = o.density_loss()
        return float((cx * o.read(0)).sum() + (cv * o.read(1)).sum() + 100 * Ld)

    rng = np.random.default_rng(5)
    eps = 1e-3
    for _ in range(16):
        d = rng.normal(size=(n, 3))
        d[nf:] = 0
        d /= np.linalg.norm(d)
        fd = (loss(v_init + eps * d) - loss(v_init - eps * d)) / (2 * eps)
        an = float((gv0 * d).sum())
        # unit direction: the error is measured against the gradient norm (a
        # directional derivative far smaller than |g| would otherwise amplify
        # the FP32 rounding of the adjoint)
        assert abs(fd - an) <= 1e-3 * np.linalg.norm(gv0[:nf]), (fd, an)
    k0 = float(p.xsph_k)
    fd = (loss(k=k0 + eps * k0) - loss(k=k0 - eps * k0)) / (2 * eps * k0)
    assert abs(fd - gk) <= 1e-3 * abs(fd), (fd, gk)


def test_torch_autograd_function():
    torch = pytest.importorskip("torch")
    from vtp.autodiff import DifferentiablePBF
    wl = scene(seed=6)
    p, ctx = setup(wl)
    nf = wl.n_fluid
    dpbf = DifferentiablePBF(ctx, STEPS, CKPT)
    v_init = ctx.read("vel")
    target = np.array([0.7, 0.2, 0.5])
    # direct C-ABI path: loss = |com - target|^2 + density loss
    ctx.rollout_fwd(STEPS, CKPT)
    x = ctx.read("pos")
    com = x[:nf].astype(np.float64).mean(0)
    _, gL = ctx.density_loss()
    gx = np.zeros((wl.n, 3), np.float32)
    gx[:nf] = 2 * (com - target) / nf
    ref_v0, _, ref_k = ctx.rollout_bwd(gx + gL, None)
    # torch path (the rollout restarts from the same initial state)
    v0 = torch.tensor(v_init, requires_grad=True)
    k = torch.tensor(float(p.xsph_k), requires_grad=True)
    xt, vt = dpbf.rollout(v0, k)
    loss = ((xt[:nf].double().mean(0) - torch.tensor(target)) ** 2).sum() + dpbf.density_loss(xt)
    loss.backward()
    np.testing.assert_allclose(v0.grad.numpy(), ref_v0, rtol=1e-4, atol=1e-7)
    assert abs(float(k.grad) - ref_k) <= 1e-4 * abs(ref_k)


def test_torch_autograd_on_device_tensors_matches_host_path():
    """The same rollout and loss with tensors on the GPU (device-pointer entry
    points, no host copies) and on the CPU (host entry points): the rollout
    and the density loss bit for bit (the same kernels on the same state);
    the gradients within 1e-5, since torch's own reductions of the loss
    terms round differently on the two devices."""
    torch = pytest.importorskip("torch")
    from vtp.autodiff import DifferentiablePBF
    wl = scene(seed=7)
    p, ctx = setup(wl)
    nf = wl.n_fluid
    dpbf = DifferentiablePBF(ctx, STEPS, CKPT)
    v_init = ctx.read("vel")
    target = torch.tensor([0.7, 0.2, 0.5], dtype=torch.float64)
    out = {}
    for dev in ("cpu", "cuda"):
        v0 = torch.tensor(v_init, device=dev, requires_grad=True)
        k = torch.tensor(float(p.xsph_k), device=dev, requires_grad=True)
        xt, vt = dpbf.rollout(v0, k)
        assert xt.device.type == dev and vt.device.type == dev
        dl = dpbf.density_loss(xt)
        loss = ((xt[:nf].double().mean(0) - target.to(dev)) ** 2).sum() + dl + (vt[:nf].double() ** 2).mean()
        loss.backward()
        out[dev] = [t.detach().cpu().numpy() for t in (xt, vt, dl, v0.grad, k.grad)]
    for name, a, b in zip(("x_T", "v_T", "density loss"), out["cpu"], out["cuda"]):
        np.testing.assert_array_equal(a, b, err_msg=name)
    np.testing.assert_allclose(out["cuda"][3], out["cpu"][3], rtol=1e-5, atol=1e-9 * np.abs(out["cpu"][3]).max())
    assert abs(float(out["cuda"][4]) - float(out["cpu"][4])) <= 1e-5 * abs(float(out["cpu"][4]))
    assert np.abs(out["cuda"][3]).max() > 0 and out["cuda"][4] != 0


def test_device_read_write_round_trip_and_errors():
    """vtp_write_device / vtp_read_device move attach-order float[n][3]
    arrays through device memory: a written velocity field reads back
    unchanged (device and host paths agree), and wrong sizes or fields are
    refused with VTP_EINVAL."""
    torch = pytest.importorskip("torch")
    wl = scene(seed=8)
    p, ctx = setup(wl)
    n = wl.n
    v = torch.randn((n, 3), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.write_device("vel", v.data_ptr())
    back = torch.empty_like(v)
    ctx.read_device("vel", back.data_ptr())
    np.testing.assert_array_equal(back.cpu().numpy(), v.cpu().numpy())
    np.testing.assert_array_equal(ctx.read("vel"), v.cpu().numpy())
    lib = vtp.load_library()
    assert lib.vtp_read_device(ctx.h, vtp._abi.FIELDS["pos"][0], back.data_ptr(), 12 * n - 12) == vtp._abi.VTP_EINVAL
    assert lib.vtp_write_device(ctx.h, vtp._abi.FIELDS["density"][0], back.data_ptr(), 12 * n) == vtp._abi.VTP_EINVAL
    assert lib.vtp_read_device(ctx.h, vtp._abi.FIELDS["pos"][0], None, 12 * n) == vtp._abi.VTP_EINVAL
    ctx.close()


def _cluster_ctx(k, K=1, xsph=0.0, seed=2):
    """k^3 jittered fluid lattice at spacing 0.018 (no walls) with N(0, 0.1^2)
    velocities, config-0 physics otherwise."""
    s = 0.018
    g = np.stack(np.meshgrid(*(np.arange(k) * s,) * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    pos = scenes.jitter(g + 0.3, s, seed=seed)
    vel = np.random.default_rng(seed).normal(0, 0.1, pos.shape).astype(np.float32)
    cfg = scenes.pbf_config(lb=(-0.5, -0.5, -0.5), size=(2, 2, 2), iterations=K, xsph=xsph)
    import json
    p = vtp.params_from_json("PBF", json.dumps(cfg))
    ctx = vtp.Context(p)
    ctx.add_particles(pos, vtp.MAT_FLUID, vel)
    ctx.initialize()
    n = pos.shape[0]
    o = pyoracle.Oracle(p, pos, vel, None, None, (np.array([0], np.uint32), np.array([n], np.uint32),
                                                   np.array([vtp.MAT_FLUID], np.int32)))
    o.initialize()
    return ctx, o, n


def test_gpu_adjoint_fully_clamped_cluster():
    """Eight particles never reach rest density: every constraint is clamped
    (C = 0, lambda = 0), so the constraint Jacobian vanishes and a velocity
    loss pulls back to the initial positions with zero gradient."""
    ctx, o, n = _cluster_ctx(2)
    cv = np.random.default_rng(1).normal(size=(n, 3))
    ctx.rollout_fwd(1, 1)
    gv0, gx0, _ = ctx.rollout_bwd(None, cv)
    assert np.all(gx0 == 0.0)
    np.testing.assert_allclose(gv0, cv, rtol=1e-6)


@pytest.mark.parametrize("k,K,xsph,steps", [(5, 1, 0.0, 1), (8, 1, 0.0, 1), (8, 4, 0.01, 1), (8, 4, 0.01, 3)])
def test_gpu_adjoint_cluster_matches_fp64(k, K, xsph, steps):
    ctx, o, n = _cluster_ctx(k, K, xsph)
    rng = np.random.default_rng(3)
    cx, cv = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    ctx.rollout_fwd(steps, 2)
    gv0, gx0, gk = ctx.rollout_bwd(cx, cv)
    o.rollout_fwd(steps)
    rv0, rx0, rk = o.rollout_bwd(cx, cv)
    assert np.linalg.norm(gx0 - rx0) <= 1e-3 * np.linalg.norm(rx0)
    assert np.linalg.norm(gv0 - rv0) <= 1e-3 * np.linalg.norm(rv0)
    if xsph:
        assert abs(gk - rk) <= 1e-3 * abs(rk)


def test_backward_graph_replay_matches_eager():
    # The first backward captures one graph per checkpoint segment and later
    # ones replay them; with stage timing on, the same work is launched
    # eagerly.  All three give the same gradients bit for bit, also after an
    # XSPH coefficient change (which must drop the captured graphs).
    wl = scene()
    p, ctx = setup(wl)
    cx, cv = loss_weights(wl)
    x0, v0 = ctx.read("pos"), ctx.read("vel")

    def grads():
        ctx.write("pos", x0)
        ctx.write("vel", v0)
        ctx.rollout_fwd(STEPS, CKPT)
        return ctx.rollout_bwd(cx, cv)

    first = grads()
    replay = grads()
    ctx.set_option("stage_timing", 1)
    eager = grads()
    ctx.set_option("stage_timing", 0)
    for a, b, c in zip(first[:2], replay[:2], eager[:2]):
        assert np.array_equal(a, b) and np.array_equal(a, c)
    assert first[2] == replay[2] == eager[2]
    ctx.set_xsph(0.02)
    changed = grads()
    ctx.set_option("stage_timing", 1)
    changed_eager = grads()
    assert not np.array_equal(changed[0], first[0])
    assert np.array_equal(changed[0], changed_eager[0]) and changed[2] == changed_eager[2]
    ctx.close()


def test_taped_last_segment_and_repeated_backward():
    # The forward tapes the ticks of its last checkpoint segment, so the first
    # backward reverses that segment without re-forwarding it; a second
    # backward of the same rollout must re-forward it (the first overwrote the
    # tapes).  A single-segment rollout is taped whole.  All give the same
    # gradients bit for bit, and the lists handed back after the taped last
    # tick give the same density loss as plain ticks.
    wl = scene()
    p, ctx = setup(wl)
    cx, cv = loss_weights(wl)
    x0, v0 = ctx.read("pos"), ctx.read("vel")

    def restart():
        ctx.write("pos", x0)
        ctx.write("vel", v0)

    restart()
    ctx.step(STEPS)
    loss_step, grad_step = ctx.density_loss()
    restart()
    ctx.rollout_fwd(STEPS, CKPT)
    loss_fwd, grad_fwd = ctx.density_loss()
    # the FP64 loss sum takes one atomic add per wavefront (order not fixed)
    assert abs(loss_fwd - loss_step) <= 1e-12 * abs(loss_step) and np.array_equal(grad_fwd, grad_step)
    first = ctx.rollout_bwd(cx, cv)
    second = ctx.rollout_bwd(cx, cv)
    restart()
    ctx.rollout_fwd(STEPS, STEPS)
    whole = ctx.rollout_bwd(cx, cv)
    whole_again = ctx.rollout_bwd(cx, cv)
    for r in (second, whole, whole_again):
        assert np.array_equal(first[0], r[0]) and np.array_equal(first[1], r[1]) and first[2] == r[2]
    ctx.close()


def test_adjoint_refuses_rollouts_with_capped_lists():
    """The reverse passes need symmetric lists (j in N(i) <=> i in N(j)),
    which a list cut by maxNeighborNum breaks.  A clump whose lists reach a
    cap of 20 (33 neighbours at rest) is refused by the forward and the
    backward with VTP_ENBRCAP; the same clump with the template cap of 60
    differentiates normally.  The context itself keeps working."""
    wl = scene()
    wl.config["PBF"]["Required"]["maxNeighborNum"] = 20
    p, ctx = setup(wl)
    with pytest.raises(vtp.VtpError, match="VTP_ENBRCAP"):
        ctx.rollout_fwd(4, 2)
    with pytest.raises(vtp.VtpError, match="VTP_ENBRCAP"):
        ctx.rollout_bwd(np.ones((wl.n, 3), np.float32), None)
    ctx.step(1)   # plain ticks are unaffected
    ctx.close()
    ok = scene()
    p, ctx = setup(ok)
    ctx.rollout_fwd(4, 2)
    gv0, _, _ = ctx.rollout_bwd(np.ones((ok.n, 3), np.float32), None)
    assert np.isfinite(gv0).all() and np.abs(gv0).max() > 0
    ctx.close()


def test_adjoint_workload_loss_16_directions():
    """The differentiable-rollout workload's own loss on a 16 x 16 x 16
    lattice (4,096 fluid particles, the 64 x 32 x 32 workload's physics, box
    and N(0, 0.1^2) initial velocities), 20 ticks in two checkpoint segments
    of 10:  L = |com_T - (0.7, 0.2, 0.5)|^2 + mean_f (rho_f / rho0 - 1)^2,
    com_T the mean fluid position after the rollout.  dL/dv0 within 1e-3
    (L2, relative) and dL/dXSPH_k within 1e-3 of the FP64 adjoint; 16 seeded
    unit directions of dL/dv0 within 1e-3 |dL/dv0| of central differences
    (eps 1e-3) of the FP64 forward, and dL/dXSPH_k within 1e-3 of a central
    difference in XSPH_k."""
    steps, ckpt = 20, 10
    target = np.array([0.7, 0.2, 0.5])
    wl = scenes.cfg4_adjoint(seed=7, n_side=(16, 16, 16))
    p, ctx = setup(wl)
    n, nf = wl.n, wl.n_fluid
    print(f"workload loss: {n} particles ({nf} fluid), {steps} ticks", flush=True)

    def com_grad(x):
        g = np.zeros((n, 3))
        g[:nf] = 2 * (x[:nf].astype(np.float64).mean(0) - target) / nf
        return g

    ctx.rollout_fwd(steps, ckpt)
    x = ctx.read("pos")
    Ld, gL = ctx.density_loss()
    Lg = float(((x[:nf].astype(np.float64).mean(0) - target) ** 2).sum() + Ld)
    gv0, _, gk = ctx.rollout_bwd(com_grad(x) + gL, np.zeros((n, 3)))
    ctx.close()
    o = pyoracle.Oracle.from_workload(p, wl)
    o.initialize()
    o.rollout_fwd(steps)
    xo = o.read(0)
    Ldo, gLo = o.density_loss()
    Lo = float(((xo[:nf].mean(0) - target) ** 2).sum() + Ldo)
    rv0, _, rk = o.rollout_bwd(com_grad(xo) + gLo, np.zeros((n, 3)))
    e = np.linalg.norm(gv0[:nf] - rv0[:nf]) / np.linalg.norm(rv0[:nf])
    print(f"loss gpu {Lg:.9g} oracle {Lo:.9g}; dL/dv0 error {e:.2e}; dL/dk gpu {gk:.6g} oracle {rk:.6g}", flush=True)
    assert abs(Lg - Lo) <= 1e-5 * abs(Lo)
    assert e <= 1e-3, e
    assert abs(gk - rk) <= 1e-3 * abs(rk), (gk, rk)
    v_init = np.concatenate([ob.vel if ob.vel is not None else np.zeros_like(ob.pos) for ob in wl.objects])

    def loss(vel0=None, k=None):
        q = pyoracle.Oracle.from_workload(p, wl)
        q.initialize()
        if vel0 is not None:
            q.write(1, vel0)
        if k is not None:
            q.set_xsph(k)
        q.step(steps)
        L, _ = q.density_loss()
        return float(((q.read(0)[:nf].mean(0) - target) ** 2).sum() + L)

    rng = np.random.default_rng(11)
    eps, gn = 1e-3, np.linalg.norm(gv0[:nf])
    worst = 0.0
    for _ in range(16):
        d = rng.normal(size=(n, 3))
        d[nf:] = 0
        d /= np.linalg.norm(d)
        fd = (loss(v_init + eps * d) - loss(v_init - eps * d)) / (2 * eps)
        an = float((gv0 * d).sum())
        worst = max(worst, abs(fd - an) / gn)
        assert abs(fd - an) <= 1e-3 * gn, (fd, an, gn)
    k0 = float(p.xsph_k)
    fdk = (loss(k=k0 + eps * k0) - loss(k=k0 - eps * k0)) / (2 * eps * k0)
    print(f"16 directions: worst |fd - adjoint| / |dL/dv0| = {worst:.2e}; dL/dk central difference {fdk:.6g}",
          flush=True)
    assert abs(fdk - gk) <= 1e-3 * abs(fdk), (fdk, gk)
