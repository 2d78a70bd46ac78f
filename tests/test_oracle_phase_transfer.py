"""Pins of the oracle's IMM phase transfer (update_phase_change_from_drift /
_from_diffuse, IMMCudaImpl.cu:532-605, and the negative-fraction loop) by
what the mathematics fixes, without restating the oracle's flux formulas:

* every pair exchange is antisymmetric (the drift term through
  alpha_i d_i + alpha_j d_j with gradW_ji = -gradW_ij, the diffusion term
  through alpha_i - alpha_j), and the drift sums to zero over the phases, so
  with equal fluid volumes the total of every phase is conserved through
  whole ticks, negative-fraction retries included (a planted
  alpha_i d_i - alpha_j d_j breaks it);
* diffusion runs down the gradient: two isolated particles with different
  fractions, equal phase densities (no drift) and no gravity end closer.
"""
import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes


def _oracle(wl):
    params = vtp.params_from_json(wl.solver, wl.config_json())
    o = pyoracle.Oracle.from_workload(params, wl)
    o.initialize()
    return o


@pytest.mark.parametrize("solver", ["IMM", "IMMCT"])
def test_phase_volumes_conserved(solver):
    wl = scenes.cfg2_immct(n_side=(12, 12, 12), box=(0.3, 0.3, 0.3), solver=solver)
    wl.config[solver]["Required"]["timeStep"] = 0.002       # retries in the negative-fraction loop
    o = _oracle(wl)
    fid = vtp._abi.FIELDS["vol_frac"][0]
    nf = wl.n_fluid
    s0 = o.read(fid)[:nf].sum(0)
    rounds = []
    for _ in range(4):
        o.step(1)
        rounds.append(o.stats()["phase_iterations"])
        s = o.read(fid)[:nf].sum(0)
        np.testing.assert_allclose(s, s0, rtol=1e-12, atol=0)
    assert max(rounds) >= 2


def test_diffusion_runs_down_the_gradient():
    cfg = scenes.immct_config(lb=(-0.05, -0.05, -0.05), size=(0.2, 0.2, 0.2), solver="IMM")
    req = cfg["IMM"]["Required"]
    req["phaseRestDensity"] = [1000.0, 1000.0, 1000.0]
    req["diffusionCoefficientCf"] = 0.01
    cfg["IMM"]["Optional"]["gravity"] = [0.0, 0.0, 0.0]
    pos = np.array([[0.05, 0.05, 0.05], [0.06, 0.05, 0.05]], np.float32)   # h = 0.02
    frac = np.array([[0.625, 0.25, 0.125], [0.125, 0.375, 0.5]], np.float32)   # exact binary fractions
    wl = scenes.Workload("pair", "IMM", cfg, [scenes.Obj(vtp.MAT_FLUID, pos, None, frac)])
    o = _oracle(wl)
    fid = vtp._abi.FIELDS["vol_frac"][0]
    a0 = o.read(fid)
    o.step(1)
    assert o.stats()["phase_iterations"] == 1   # no fraction went negative: nothing excluded
    a1 = o.read(fid)
    d0, d1 = a0[0] - a0[1], a1[0] - a1[1]
    assert np.all(np.abs(d1) < np.abs(d0)) and np.all(np.sign(d1) == np.sign(d0))
    np.testing.assert_allclose(a1.sum(0), a0.sum(0), rtol=1e-12)
