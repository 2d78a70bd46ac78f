"""Pins of the oracle's IMM-CT conformation update (update_ct_cuda,
IMMCTCudaImpl.cu:716-744) by closed forms of the ODE it integrates,

    Q' = Q L + L^T Q - (Q - I)/lambda - beta (Q - I) Q,

and of the viscosity law, without restating the oracle's step:

* simple shear (L_xy = g, beta = 0): the explicit-Euler iteration converges to
  the steady state Q_xx = 1, Q_xy = Q_yx = lambda g, Q_yy = 1 + 2 (lambda g)^2,
  Q_zz = 1 (a transposed Q L^T + L Q would put the 2 (lambda g)^2 on Q_xx);
* no flow, beta > 0, Q = diag(1 + u0, 1, 1): u' = -a u - beta u^2 with
  a = 1/lambda + beta, whose solution is u0 a e^{-a t} / (a + beta u0 (1 - e^{-a t}));
* viscosity: the shear rate is sqrt(tr(D)^2 / 2) (the reference squares the
  trace), so a traceless flow of any strength leaves mu = min(base kappa, max);
  10 * rate = 1 halves the thickening and 10 * rate = 2 divides it by 33;
  stress = mu (Q_new - I).
"""
import math

import numpy as np

import pyoracle

I3 = np.eye(3)


def _step(Q, L, dt, relax=0.5, beta=0.0, kappa=2.0, base=1.0, vmax=1.5, a1=0.6, th=0.5):
    return pyoracle.ct_update(Q, L, dt, relax, beta, kappa, base, vmax, a1, th)


def test_simple_shear_steady_state():
    lam, g, dt = 0.5, 1.3, 1e-3
    L = np.zeros((3, 3))
    L[0, 1] = g
    Q = I3.copy()
    for _ in range(int(60 * lam / dt)):
        Q, _, _ = _step(Q, L, dt, relax=lam)
    wi = lam * g
    expect = np.array([[1.0, wi, 0.0], [wi, 1.0 + 2.0 * wi * wi, 0.0], [0.0, 0.0, 1.0]])
    np.testing.assert_allclose(Q, expect, rtol=0, atol=1e-9)


def test_thinning_term_matches_the_logistic_solution():
    lam, beta, u0, dt, T = 0.4, 3.0, 0.8, 1e-5, 0.3
    Q = np.diag([1.0 + u0, 1.0, 1.0])
    L = np.zeros((3, 3))
    for _ in range(int(round(T / dt))):
        Q, _, _ = _step(Q, L, dt, relax=lam, beta=beta)
    a = 1.0 / lam + beta
    e = math.exp(-a * T)
    u = u0 * a * e / (a + beta * u0 * (1.0 - e))
    assert abs((Q[0, 0] - 1.0) - u) <= 2e-4 * u
    np.testing.assert_allclose(Q - np.diag(np.diag(Q)), 0.0, atol=0)
    assert Q[1, 1] == 1.0 and Q[2, 2] == 1.0


def test_viscosity_law_and_stress():
    base, vmax, kappa, th = 1.0, 1.5, 2.0, 0.5
    Q = I3 + 0.1 * np.arange(9).reshape(3, 3) / 9.0
    # traceless flows, however strong, do not thin (the trace is squared)
    for L in (np.zeros((3, 3)), np.array([[0, 50.0, 0], [0, 0, 0], [0, 0, 0]]), np.diag([7.0, -7.0, 0.0])):
        _, mu, _ = _step(Q, L, 1e-4, kappa=kappa, base=base, vmax=vmax, a1=0.6, th=th)
        assert mu == min(base * kappa, vmax)
    # 10 * rate = 1: (10 rate)^5 = 1 halves the thickening
    c = 0.1 * math.sqrt(2.0) / 3.0           # tr(D) = 3c, rate = 3c / sqrt(2) = 0.1
    _, mu, _ = _step(Q, c * I3, 1e-4, kappa=kappa, base=base, vmax=vmax, a1=0.6, th=th)
    assert abs(mu - (base + 0.5 * (vmax - base))) <= 1e-12
    # 10 * rate = 2: thickening divided by 1 + 2^5
    _, mu, _ = _step(Q, 2.0 * c * I3, 1e-4, kappa=kappa, base=base, vmax=vmax, a1=0.6, th=th)
    assert abs(mu - (base + (vmax - base) / 33.0)) <= 1e-12
    # a fast compression removes it
    _, mu, _ = _step(Q, 10.0 * I3, 1e-4, kappa=kappa, base=base, vmax=vmax, a1=0.6, th=th)
    assert abs(mu - base) <= 1e-9
    # stress = mu (Q_new - I)
    rng = np.random.default_rng(2)
    L = rng.normal(size=(3, 3))
    Qn, mu, st = _step(Q, L, 1e-3, kappa=kappa, base=base, vmax=vmax, a1=0.6, th=th)
    np.testing.assert_allclose(st, mu * (Qn - I3), rtol=0, atol=1e-15)
