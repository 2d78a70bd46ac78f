"""The FP64 PBF adjoint of the oracle against central finite differences of
the oracle's own forward rollout (pins every transposed Jacobian: XSPH,
constraint iterations through lambda, C, S and the density clamp, the
Hessian of W, the predict step and the density loss)."""
import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes


def small_scene(seed=0, side=6, box=0.25):
    wl = scenes.cfg0_pbf(n_side=(side, side, side))
    wl.config["PBF"]["Required"]["simSpaceSize"] = [0.5, 0.5, 0.5]
    wl.objects[1].pos = scenes.particle_box(0.01, (0, 0, 0), (box, box, box), 2)
    rng = np.random.default_rng(seed)
    wl.objects[0].vel = rng.normal(0, 0.1, (wl.n_fluid, 3)).astype(np.float32)
    return wl


def test_oracle_adjoint_matches_central_differences():
    wl = small_scene()
    p = vtp.params_from_json("PBF", wl.config_json())
    n, nf = wl.n, wl.n_fluid
    rng = np.random.default_rng(1)
    cx = rng.normal(size=(n, 3))
    cv = rng.normal(size=(n, 3))
    cx[nf:] = 0
    cv[nf:] = 0
    steps = 3
    v_init = np.concatenate([wl.objects[0].vel, np.zeros((n - nf, 3))])

    def loss(vel0=None, k=None):
        o = pyoracle.Oracle.from_workload(p, wl)
        o.initialize()
        if vel0 is not None:
            o.write(1, vel0)
        if k is not None:
            o.set_xsph(k)
        o.step(steps)
        L, _ = o.density_loss()
        return float((cx * o.read(0)).sum() + (cv * o.read(1)).sum() + 100 * L)

    o = pyoracle.Oracle.from_workload(p, wl)
    o.initialize()
    o.rollout_fwd(steps)
    L, gL = o.density_loss()
    gv0, _gx0, gk = o.rollout_bwd(cx + 100 * gL, cv)
    eps = 1e-6
    for _ in range(4):
        d = rng.normal(size=(n, 3))
        d[nf:] = 0
        fd = (loss(v_init + eps * d) - loss(v_init - eps * d)) / (2 * eps)
        an = float((gv0 * d).sum())
        assert abs(fd - an) <= 1e-6 * abs(fd), (fd, an)
    k0 = float(p.xsph_k)
    fd = (loss(k=k0 + eps) - loss(k=k0 - eps)) / (2 * eps)
    assert abs(fd - gk) <= 1e-6 * abs(fd), (fd, gk)


def test_oracle_density_loss_gradient():
    wl = small_scene(seed=2)
    p = vtp.params_from_json("PBF", wl.config_json())
    o = pyoracle.Oracle.from_workload(p, wl)
    o.initialize()
    o.step(2)
    L, g = o.density_loss()
    x = o.read(0)
    rng = np.random.default_rng(3)
    d = rng.normal(size=x.shape)
    d[wl.n_fluid:] = 0
    eps = 1e-7
    o.write(0, x + eps * d)
    Lp, _ = o.density_loss()
    o.write(0, x - eps * d)
    Lm, _ = o.density_loss()
    fd = (Lp - Lm) / (2 * eps)
    assert abs(fd - (g * d).sum()) <= 1e-6 * abs(fd)
