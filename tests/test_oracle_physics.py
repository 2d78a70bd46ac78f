"""Pins of the FP64 oracle's solver steps by what the mathematics fixes,
independently of the oracle's own formulas:

PBF    free fall of an isolated particle in closed form; density and lambda of
       an interior particle of a compressed lattice from a direct FP64 sum over
       lattice offsets; exact conservation of total momentum (sum v) of an
       isolated fluid cluster with gravity off (the constraint and XSPH passes
       are pairwise antisymmetric).
DFSPH  compression ratio of an interior lattice particle from the lattice
       sum; a uniformly translating, uncompressed cluster keeps its velocity
       exactly (viscosity of a uniform field is zero, no pressure); total
       momentum of an isolated cluster with pressure, viscosity and surface
       tension is conserved.
IMM    phase fractions sum to one and stay non-negative; a uniform fraction
       field with identical phases stays uniform.
IMM-CT a rigidly translating cluster has zero velocity gradient and keeps
       Q = I.

A dropped term, a flipped sign or a wrong neighbour weight breaks one of
these; restating the oracle's formulas here would not.
"""
import json
import math

import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes

F_POS, F_VEL, F_MASS, F_DENS, F_LAMBDA, F_CR, F_VOLFRAC, F_Q, F_L = 0, 1, 3, 5, 6, 9, 18, 22, 23


def _no_gravity(cfg, domain):
    cfg[domain]["Optional"] = {"enable": True, "gravity": [0.0, 0.0, 0.0]}
    return cfg


def _oracle(solver, cfg, pos, vel=None, frac=None):
    p = vtp.params_from_json(solver, json.dumps(cfg))
    n = pos.shape[0]
    o = pyoracle.Oracle(p, pos, vel, frac, None, (np.array([0], np.uint32), np.array([n], np.uint32),
                                                  np.array([vtp.MAT_FLUID], np.int32)))
    o.initialize()
    return o, p


def _cubic(h, q):
    s = 8.0 / (3.14159265 * h ** 3)
    if q <= 0.5:
        return s * (6 * q ** 3 - 6 * q ** 2 + 1)
    if q <= 1:
        return s * 2 * (1 - q) ** 3
    return 0.0


def _cubic_grad_norm(h, q):
    s = 8.0 / (3.14159265 * h ** 3)
    if q < 1e-6 or q > 1:
        return 0.0
    return s / h * abs(6 * (3 * q * q - 2 * q) if q <= 0.5 else -6 * (1 - q) ** 2)


def _lattice(spacing, n_side, origin=0.1):
    a = origin + spacing * np.arange(n_side)
    Z, Y, X = np.meshgrid(a, a, a, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)


def _lattice_offsets(spacing, h):
    m = int(math.ceil(h / spacing)) + 1
    out = []
    for i in range(-m, m + 1):
        for j in range(-m, m + 1):
            for k in range(-m, m + 1):
                d = spacing * math.sqrt(i * i + j * j + k * k)
                if d <= h:
                    out.append(d)
    return out


def _interior(pos, h):
    lo, hi = pos.min(0), pos.max(0)
    return np.all((pos > lo + h * 1.01) & (pos < hi - h * 1.01), axis=1)


# ------------------------------------------------------------------ PBF
def test_pbf_free_fall_closed_form():
    cfg = scenes.pbf_config(lb=(-1, -1, -1), size=(2, 2, 2))
    x0 = np.array([[0.1, 0.5, -0.2]], np.float32)
    v0 = np.array([[0.3, 0.2, -0.1]], np.float32)
    o, p = _oracle("PBF", cfg, x0, v0)
    steps = 25
    o.step(steps)
    dt = float(p.time_step)
    g = np.array(p.gravity, np.float64)
    v = v0[0].astype(np.float64)
    x = x0[0].astype(np.float64)
    for _ in range(steps):
        v = v + g * dt
        x = x + v * dt
    np.testing.assert_allclose(o.read(F_VEL)[0], v, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(o.read(F_POS)[0], x, rtol=1e-12, atol=1e-14)


def test_pbf_compressed_lattice_density_and_lambda():
    r = 0.01
    h, s = 4 * r, 0.9 * 2 * r
    cfg = _no_gravity(scenes.pbf_config(radius=r, lb=(0, 0, 0), size=(0.5, 0.5, 0.5), iterations=1, xsph=0.0),
                      "PBF")
    pos = _lattice(s, 12)
    o, p = _oracle("PBF", cfg, pos)
    o.step(1)
    rho0 = float(p.f_rest_density)
    m = rho0 * (2 * r) ** 3
    offs = _lattice_offsets(s, h)
    rho = m * sum(_cubic(h, d / h) for d in offs)
    assert rho > 1.05 * rho0
    g2 = sum(_cubic_grad_norm(h, d / h) ** 2 for d in offs)
    lam = -(rho / rho0 - 1.0) / (g2 / rho0 ** 2 + 1e-6)   # sum_j gradW vanishes by symmetry
    inner = _interior(pos, h)
    assert inner.sum() >= 64
    np.testing.assert_allclose(o.read(F_DENS)[inner], rho, rtol=2e-7)
    np.testing.assert_allclose(o.read(F_LAMBDA)[inner], lam, rtol=2e-6)


def test_pbf_momentum_of_isolated_cluster_is_conserved():
    r = 0.01
    cfg = _no_gravity(scenes.pbf_config(radius=r, lb=(0, 0, 0), size=(0.5, 0.5, 0.5), iterations=4, xsph=0.05),
                      "PBF")
    pos = scenes.jitter(_lattice(0.9 * 2 * r, 9), 0.018, seed=3)
    rng = np.random.default_rng(4)
    vel = rng.normal(0, 0.2, pos.shape).astype(np.float32)
    o, p = _oracle("PBF", cfg, pos, vel)
    p0 = o.read(F_VEL).sum(0)
    x0 = o.read(F_POS).sum(0)
    steps = 3
    o.step(steps)
    v = o.read(F_VEL)
    assert np.abs(v - vel).max() > 1e-3                     # the constraints did move particles
    np.testing.assert_allclose(v.sum(0), p0, atol=1e-11 * np.abs(vel).sum())
    np.testing.assert_allclose(o.read(F_POS).sum(0), x0 + steps * float(p.time_step) * p0, atol=1e-11)


# ------------------------------------------------------------------ DFSPH
def _dfsph_cfg(r, tension=0.0, vis=1e-3):
    return _no_gravity(scenes.dfsph_config(radius=r, lb=(0, 0, 0), size=(0.5, 0.5, 0.5), dt=0.002, vis=vis,
                                           tension=tension), "DFSPH")


def test_dfsph_compression_ratio_of_lattice():
    r = 0.01
    h, s = 4 * r, 0.9 * 2 * r
    pos = _lattice(s, 12)
    o, p = _oracle("DFSPH", _dfsph_cfg(r), pos)
    o.step(1)
    V0 = (2 * r) ** 3
    cr = V0 * sum(_cubic(h, d / h) for d in _lattice_offsets(s, h))
    inner = _interior(pos, h)
    np.testing.assert_allclose(o.read(F_CR)[inner], cr, rtol=2e-7)


def test_dfsph_uniform_translation_is_unchanged():
    r = 0.01
    pos = _lattice(1.1 * 2 * r, 8)                            # expanded: no compression, no pressure
    u = np.array([0.3, -0.1, 0.2], np.float32)
    vel = np.tile(u, (pos.shape[0], 1))
    o, p = _oracle("DFSPH", _dfsph_cfg(r), pos, vel)
    steps = 3
    o.step(steps)
    np.testing.assert_allclose(o.read(F_VEL), np.tile(u.astype(np.float64), (pos.shape[0], 1)), atol=1e-12)
    np.testing.assert_allclose(o.read(F_POS), pos + steps * float(p.time_step) * u.astype(np.float64), atol=1e-12)


def test_dfsph_momentum_of_isolated_cluster_is_conserved():
    r = 0.01
    pos = scenes.jitter(_lattice(0.9 * 2 * r, 9), 0.018, seed=5)
    rng = np.random.default_rng(6)
    vel = rng.normal(0, 0.2, pos.shape).astype(np.float32)
    o, p = _oracle("DFSPH", _dfsph_cfg(r, tension=0.05, vis=0.01), pos, vel)
    m = o.read(F_MASS)
    p0 = (m[:, None] * o.read(F_VEL)).sum(0)
    o.step(3)
    st = o.stats()
    assert st["incomp_iterations"] >= 1
    v = o.read(F_VEL)
    assert np.abs(v - vel).max() > 1e-3
    np.testing.assert_allclose((m[:, None] * v).sum(0), p0, atol=1e-11 * np.abs(m[:, None] * vel).sum())


# ------------------------------------------------------------------ IMM / IMM-CT
def _imm_cfg(r, solver, rho, mu):
    cfg = scenes.immct_config(radius=r, lb=(0, 0, 0), size=(0.5, 0.5, 0.5), solver=solver)
    req = cfg[solver]["Required"]
    req["phaseRestDensity"] = list(rho)
    if solver == "IMM":
        req["phaseRestViscosity"] = list(mu)
    return _no_gravity(cfg, solver)


def test_imm_fractions_stay_normalised_and_non_negative():
    r = 0.01
    pos = scenes.jitter(_lattice(0.9 * 2 * r, 8), 0.018, seed=7)
    rng = np.random.default_rng(8)
    frac = rng.dirichlet([0.3, 0.3, 0.3], pos.shape[0]).astype(np.float32)
    vel = rng.normal(0, 0.1, pos.shape).astype(np.float32)
    o, p = _oracle("IMM", _imm_cfg(r, "IMM", (1000, 800, 1200), (1e-3, 5e-3, 1e-2)), pos, vel, frac)
    o.step(3)
    a = o.read(F_VOLFRAC)
    assert np.abs(a - frac).max() > 1e-4                       # drift and diffusion did move mass
    np.testing.assert_allclose(a.sum(1), 1.0, atol=1e-12)
    assert a.min() >= 0.0


def test_imm_identical_phases_keep_uniform_fractions():
    r = 0.01
    pos = scenes.jitter(_lattice(0.9 * 2 * r, 8), 0.018, seed=9)
    frac = np.tile(np.array([0.2, 0.5, 0.3], np.float32), (pos.shape[0], 1))
    o, p = _oracle("IMM", _imm_cfg(r, "IMM", (1000, 1000, 1000), (1e-3, 1e-3, 1e-3)), pos, None, frac)
    o.step(3)
    np.testing.assert_allclose(o.read(F_VOLFRAC), frac.astype(np.float64), atol=1e-6)


def test_immct_rigid_translation_keeps_identity_conformation():
    r = 0.01
    pos = _lattice(1.1 * 2 * r, 7)
    u = np.array([0.2, 0.1, -0.3], np.float32)
    frac = np.tile(np.array([0.3, 0.6, 0.1], np.float32), (pos.shape[0], 1))
    cfg = _imm_cfg(r, "IMMCT", (1000, 1000, 1000), None)
    o, p = _oracle("IMMCT", cfg, pos, np.tile(u, (pos.shape[0], 1)), frac)
    o.step(3)
    assert np.abs(o.read(F_L)).max() <= 1e-9
    np.testing.assert_allclose(o.read(F_Q), np.tile(np.eye(3).ravel(), (pos.shape[0], 1)), atol=1e-12)
