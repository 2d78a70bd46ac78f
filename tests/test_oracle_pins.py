"""Closed-form pins of the oracle's solver passes on exact lattices.

Every lattice here sits on multiples of a power of two (origin 0.25, spacing
2^-6 or 3*2^-7) with particle radius 5*2^-9, so positions, neighbour offsets,
the smoothing radius h = 4r = 2.5 * 2^-6 and the rest volume (2r)^3 are exact
in FP32 and the oracle's FP64 sums see exactly the offsets the tests sum over.
The expected values are lattice sums of the cubic spline (its closed form is
pinned in test_oracle_kernels.py) taken over the integer offsets o with
|o| s <= h, written from the reference's formulas and the symmetry of the
lattice, not from the oracle's loops:

* DFSPH (DFSPHCudaImpl.cu:85-220): the compression ratio, alpha and the
  first iteration's kappa of an interior particle of a compressed block; on a
  z-layered slab the per-layer kappa (alpha's |sum V gradW|^2 / m term is
  non-zero near the faces) and the v* change of the top face's particle.
* PBF (PBFCudaImpl.cu:70-194): per-layer density, lambda and the position
  correction of a slab's top layer; XSPH of the shear field v = (y^2, 0, 0),
  whose interior value is k V0 sum_o (o_y s)^2 W.
* IMM (IMMCudaImpl.cu:532-605): the drift flux of a linear drift field d = G x
  changes the fraction by dt a V0 tr(G) S_f and the diffusion of a quadratic
  field a = a0 + c x^2 by -dt Cf V0 c S_f, where S_f = sum_o f(|o|s) (o_x s)^2
  and f = W'(r)/r.
* IMM-CT (IMMCTCudaImpl.cu:687-821): the velocity gradient of a linear field
  v = A x is L = -V0 S_f A (a transposed outer product gives A^T); Cd of a
  uniform fraction a is Cd0 + (1 - Cd0) a V0 sum_{o != 0} W; the rheological
  factors at special fractions; the polymer-stress acceleration of a slab's
  face particle under a uniform stress S is (2 a_p / rho) S sum_j gradW_ij.

A dropped factor (dt, V, rho0, 1/m, the 2 of symmetric sums), a flipped sign
or a transposed operand in the oracle fails one of these.
"""
import json
import math

import numpy as np
import pytest

import pyoracle
import vtp
from vtp import scenes

PI = 3.14159265
R = 5.0 * 2.0 ** -9            # particle radius
H = 4.0 * R                    # smoothing radius, 2.5 * 2^-6
V0 = (2.0 * R) ** 3            # rest volume
S_C = 2.0 ** -6                # compressed spacing, 0.8 * 2r
S_E = 3.0 * 2.0 ** -7          # expanded spacing, 1.2 * 2r: compression ratio < 1, no pressure
ORIGIN = 0.25

F_POS, F_VEL, F_DENS, F_LAMBDA, F_CR, F_ALPHA, F_DELTA, F_KDIV, F_KINC, F_VADV = 0, 1, 5, 6, 9, 10, 11, 12, 13, 14
F_VIS, F_FRAC, F_VPH, F_CD, F_Q, F_L, F_STRESS = 17, 18, 19, 21, 22, 23, 24


# ------------------------------------------------------------ kernel closed form
def w(r):
    q = r / H
    s = 8.0 / (PI * H ** 3)
    if q <= 0.5:
        return s * (6 * q ** 3 - 6 * q ** 2 + 1)
    if q <= 1.0:
        return s * 2 * (1 - q) ** 3
    return 0.0


def f(r):
    """W'(r) / r, so that gradW(rvec) = f(|rvec|) rvec (zero at q < 1e-6 and q > 1)."""
    q = r / H
    if q < 1e-6 or q > 1.0:
        return 0.0
    s = 8.0 / (PI * H ** 3)
    dwdq = s * (18 * q * q - 12 * q) if q <= 0.5 else -6 * s * (1 - q) ** 2
    return dwdq / H / r


def grad(rvec):
    rvec = np.asarray(rvec, np.float64)
    return f(float(np.linalg.norm(rvec))) * rvec


def offsets(s):
    m = int(math.ceil(H / s)) + 1
    out = []
    for a in range(-m, m + 1):
        for b in range(-m, m + 1):
            for c in range(-m, m + 1):
                if math.sqrt(a * a + b * b + c * c) * s <= H:
                    out.append((a, b, c))
    return out


def s_f(s):
    """sum_o f(|o| s) (o_x s)^2 over the full lattice stencil (negative)."""
    return sum(f(math.sqrt(a * a + b * b + c * c) * s) * (a * s) ** 2 for a, b, c in offsets(s))


# ------------------------------------------------------------ lattices and oracle runs
def box_lattice(nx, ny, nz, s):
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1)
    pos = (ORIGIN + s * idx).astype(np.float32)
    assert np.array_equal(pos.astype(np.float64), ORIGIN + s * idx.astype(np.float64))   # exact in FP32
    return pos, idx


def at(idx, i, j, k):
    return int(np.flatnonzero((idx == (i, j, k)).all(1))[0])


def _no_gravity(cfg, dom):
    cfg[dom]["Optional"] = {"enable": True, "gravity": [0.0, 0.0, 0.0]}
    cfg[dom]["Required"]["maxNeighborNum"] = 100      # 81 neighbours at 0.8 spacing: no cap
    return cfg


def run(solver, cfg, pos, vel=None, frac=None, before=None, stop=0):
    p = vtp.params_from_json(solver, json.dumps(cfg))
    n = pos.shape[0]
    o = pyoracle.Oracle(p, pos, vel, frac, None, (np.array([0], np.uint32), np.array([n], np.uint32),
                                                  np.array([vtp.MAT_FLUID], np.int32)))
    o.initialize()
    if before:
        before(o)
    o.set_stop(stop)
    o.step(1)
    return o, p


def dfsph_cfg(dt=0.002):
    cfg = scenes.dfsph_config(radius=R, lb=(0, 0, 0), size=(1, 1, 1), dt=dt, vis=0.0, div_thr=10.0, inc_thr=10.0,
                              tension=0.0)
    return _no_gravity(cfg, "DFSPH")


# --------------------------------------------------------------------------- DFSPH
def test_dfsph_interior_alpha_and_first_kappa_closed_form():
    """Interior particle of a compressed 15^3 block, v = 0, thresholds above
    the first residual (one iteration per solve).  Its 2-layer neighbourhood
    and theirs are complete, so kappa is uniform there, the v* update
    vanishes by symmetry and both solves see the same delta."""
    pos, idx = box_lattice(15, 15, 15, S_C)
    o, p = run("DFSPH", dfsph_cfg(), pos)
    dt, rho0 = float(np.float32(p.time_step)), float(p.f_rest_density)
    offs = offsets(S_C)
    assert len(offs) == 81
    cr = V0 * sum(w(S_C * math.sqrt(a * a + b * b + c * c)) for a, b, c in offs)
    g2 = sum(float(np.dot(grad(np.array(v) * S_C), grad(np.array(v) * S_C))) for v in offs)
    alpha = 1e-6 + (V0 / rho0) * g2              # V_j^2 / m_j sum |gradW|^2 with m = rho0 V0
    kappa = (cr - 1.0) / alpha / dt ** 2 / V0
    assert cr > 1.5
    c = at(idx, 7, 7, 7)
    assert o.stats()["div_iterations"] == 1 and o.stats()["incomp_iterations"] == 1
    assert o.read(F_CR)[c] == pytest.approx(cr, rel=1e-13)
    assert o.read(F_ALPHA)[c] == pytest.approx(alpha, rel=1e-13)
    assert o.read(F_KDIV)[c] == pytest.approx(kappa, rel=1e-12)
    assert o.read(F_KINC)[c] == pytest.approx(kappa, rel=1e-12)
    assert o.read(F_DELTA)[c] == pytest.approx(cr - 1.0, rel=1e-12)
    np.testing.assert_allclose(o.read(F_VEL)[c], 0.0, atol=1e-12)   # symmetric sums cancel to rounding


def _slab_layer_sums(nz, l):
    """Neighbour offsets of a particle of layer l (0..nz-1) in an unbounded-in-xy slab."""
    return [o for o in offsets(S_C) if 0 <= l + o[2] < nz]


def _dfsph_layer(nz, l, rho0, dt):
    offs = _slab_layer_sums(nz, l)
    cr = V0 * sum(w(S_C * math.sqrt(a * a + b * b + c * c)) for a, b, c in offs)
    a1 = V0 * sum(grad(-np.array(v, np.float64) * S_C) for v in offs)
    m = rho0 * V0
    a2 = 1e-6 + (V0 * V0 / m) * sum(float(np.dot(grad(np.array(v) * S_C), grad(np.array(v) * S_C))) for v in offs)
    alpha = max(float(np.dot(a1, a1)) / m + a2, 1e-6)
    kappa = max(cr - 1.0, 0.0) / alpha / dt ** 2 / V0
    return cr, alpha, kappa, a1


def test_dfsph_slab_face_kappa_and_vstar_closed_form():
    """Slab of 13 x 13 x 5 layers, v = 0, one divergence iteration (the tick
    stops after that solve).  Per layer, at the centre column: cr, alpha (with
    its |sum_j V_j gradW_ij|^2 / m_i term, which is non-zero in the layers
    near a face) and kappa_div; the top face particle's v* is
    -dt V0^2/m sum_j (kappa_i + kappa_j) gradW_ij over the layered stencil."""
    nz = 5
    pos, idx = box_lattice(13, 13, nz, S_C)
    o, p = run("DFSPH", dfsph_cfg(), pos, stop=1)
    dt, rho0 = float(np.float32(p.time_step)), float(p.f_rest_density)
    layers = [_dfsph_layer(nz, l, rho0, dt) for l in range(nz)]
    cr_o, al_o, kd_o = o.read(F_CR), o.read(F_ALPHA), o.read(F_KDIV)
    for l, (cr, alpha, kappa, a1) in enumerate(layers):
        c = at(idx, 6, 6, l)
        assert cr_o[c] == pytest.approx(cr, rel=1e-13)
        assert al_o[c] == pytest.approx(alpha, rel=1e-12)
        assert kd_o[c] == pytest.approx(kappa, rel=1e-12)
    assert abs(layers[0][3][2]) > 0 and layers[0][3][2] == pytest.approx(-layers[-1][3][2], rel=1e-12)
    top = nz - 1
    m = rho0 * V0
    dv = np.zeros(3)
    for v in _slab_layer_sums(nz, top):
        kj = layers[top + v[2]][2]
        dv -= dt * V0 * V0 / m * (layers[top][2] + kj) * grad(-np.array(v, np.float64) * S_C)
    assert abs(dv[2]) > 1e-3 and dv[2] > 0          # the compressed slab pushes its face outwards
    got = o.read(F_VADV)[at(idx, 6, 6, top)]
    np.testing.assert_allclose(got, dv, rtol=1e-11, atol=1e-12 * abs(dv[2]))


# --------------------------------------------------------------------------- PBF
def pbf_cfg(iterations, xsph, dt=0.004):
    cfg = scenes.pbf_config(radius=R, lb=(0, 0, 0), size=(1, 1, 1), dt=dt, iterations=iterations, xsph=xsph)
    return _no_gravity(cfg, "PBF")


def test_pbf_slab_top_layer_correction_closed_form():
    """One constraint iteration on a 13 x 13 x 5 slab at rest, XSPH off: the
    final position of a centre-column particle is x + dx.  Per layer l the
    density m sum W, C = max(rho, rho0)/rho0 - 1, lambda = -C/(S + 1e-6) with
    S = (sum |gradW|^2 + |sum gradW|^2)/rho0^2; the top particle moves by
    (1/rho0) sum_j (lambda_i + lambda_j) gradW_ij."""
    nz = 5
    pos, idx = box_lattice(13, 13, nz, S_C)
    o, p = run("PBF", pbf_cfg(1, 0.0), pos)
    rho0 = float(p.f_rest_density)
    m = rho0 * V0
    lam = []
    for l in range(nz):
        offs = _slab_layer_sums(nz, l)
        rho = m * sum(w(S_C * math.sqrt(a * a + b * b + c * c)) for a, b, c in offs)
        C = max(rho, rho0) / rho0 - 1.0
        gs = sum(grad(-np.array(v, np.float64) * S_C) for v in offs)
        g2 = sum(float(np.dot(grad(np.array(v) * S_C), grad(np.array(v) * S_C))) for v in offs)
        lam.append(-C / ((g2 + float(np.dot(gs, gs))) / rho0 ** 2 + 1e-6))
        c = at(idx, 6, 6, l)
        assert o.read(F_DENS)[c] == pytest.approx(max(rho, rho0), rel=1e-13)
        assert o.read(F_LAMBDA)[c] == pytest.approx(lam[-1], rel=1e-12)
    top = nz - 1
    dx = np.zeros(3)
    for v in _slab_layer_sums(nz, top):
        dx += (lam[top] + lam[top + v[2]]) * grad(-np.array(v, np.float64) * S_C)
    dx /= rho0
    assert dx[2] > 1e-5                              # the compressed slab pushes its face outwards
    c = at(idx, 6, 6, top)
    got = o.read(F_POS)[c] - pos[c].astype(np.float64)
    np.testing.assert_allclose(got, dx, rtol=1e-9, atol=1e-9 * abs(dx[2]))


def test_pbf_xsph_of_quadratic_shear_closed_form():
    """No constraint iterations, v = (y^2, 0, 0) (exact in FP32 on this
    lattice), dt 1e-5 so that the predicted positions stay a lattice to
    O(dt^2).  The XSPH pass adds k sum_j V_j (v_j - v_i) W_ij to v; at an
    interior particle the linear part of v_j - v_i = 2 y_i dy + dy^2 sums to
    zero, leaving k V0 sum_o (o_y s)^2 W(|o| s)."""
    k = 0.01
    pos, idx = box_lattice(11, 11, 11, S_C)
    vel = np.zeros_like(pos)
    vel[:, 0] = pos[:, 1].astype(np.float64) ** 2
    assert np.array_equal(vel[:, 0].astype(np.float64), pos[:, 1].astype(np.float64) ** 2)
    o, p = run("PBF", pbf_cfg(0, k, dt=1e-5), pos, vel)
    dv = k * V0 * sum((b * S_C) ** 2 * w(S_C * math.sqrt(a * a + b * b + cc * cc)) for a, b, cc in offsets(S_C))
    c = at(idx, 5, 5, 5)
    got = o.read(F_VEL)[c] - vel[c].astype(np.float64)
    assert got[0] == pytest.approx(dv, rel=1e-8)
    assert abs(got[1]) < 1e-15 and abs(got[2]) < 1e-15


# --------------------------------------------------------------------------- IMM
def imm_cfg(cf, dt=0.002):
    cfg = scenes.immct_config(radius=R, lb=(0, 0, 0), size=(1, 1, 1), dt=dt, solver="IMM")
    req = cfg["IMM"]["Required"]
    req.update({"phaseRestDensity": [1000.0, 1000.0], "phaseRestViscosity": [0.0, 0.0],
                "phaseRestColor": [[1, 0, 0], [0, 0, 1]], "diffusionCoefficientCf": cf,
                "divFreeThreshold": 10.0, "incompThreshold": 10.0})
    return _no_gravity(cfg, "IMM")


def test_imm_drift_flux_of_linear_drift_closed_form():
    """Fractions (1/2, 1/2) everywhere, phase velocities u_0 = G (x - x_c),
    u_1 = -u_0 (mixture velocity 0), no diffusion, on an uncompressed block
    (no pressure: nothing moves).  At the centre of an 11^3 block the drift
    flux changes phase 0's fraction by dt a V0 tr(G) S_f (phase 1's by the
    opposite)."""
    pos, idx = box_lattice(11, 11, 11, S_E)
    n = pos.shape[0]
    c = at(idx, 5, 5, 5)
    G = np.array([[0.3, 0.2, 0.0], [0.0, 0.5, -0.1], [0.4, 0.0, 0.2]])
    d = (pos.astype(np.float64) - pos[c].astype(np.float64)) @ G.T
    frac = np.full((n, 2), 0.5, np.float32)

    def phase_velocities(orc):
        u = np.zeros((n, 2, 3))
        u[:, 0] = d
        u[:, 1] = -d
        orc.write(F_VPH, u)

    o, p = run("IMM", imm_cfg(0.0), pos, None, frac, before=phase_velocities)
    dt = float(np.float32(p.time_step))
    change = dt * 0.5 * V0 * np.trace(G) * s_f(S_E)
    assert change < -1e-4
    a = o.read(F_FRAC)[c]
    assert a[0] - 0.5 == pytest.approx(change, rel=1e-9)
    assert a[1] - 0.5 == pytest.approx(-change, rel=1e-9)


def test_imm_diffusion_of_quadratic_fraction_closed_form():
    """Fractions a_0 = 1/2 + c |x - x_c|_x^2 (x only), a_1 = 1 - a_0, no drift.
    At the centre the diffusion flux dt Cf sum_j (a_i - a_j) V_j f_ij equals
    -dt Cf V0 c S_f: mass flows into the minimum.  Uncompressed block: no
    pressure, nothing moves."""
    pos, idx = box_lattice(11, 11, 11, S_E)
    n = pos.shape[0]
    c = at(idx, 5, 5, 5)
    cq = 4.0
    dx = pos[:, 0].astype(np.float64) - float(pos[c, 0])
    frac = np.empty((n, 2), np.float32)
    frac[:, 0] = 0.5 + cq * dx ** 2
    frac[:, 1] = 1.0 - frac[:, 0]
    assert np.array_equal(frac[:, 0].astype(np.float64), 0.5 + cq * dx ** 2)   # exact in FP32
    cf = float(np.float32(0.1))   # the coefficient as the solver holds it
    o, p = run("IMM", imm_cfg(cf), pos, None, frac)
    dt = float(np.float32(p.time_step))
    change = -dt * cf * V0 * cq * s_f(S_E)
    assert change > 1e-5
    a = o.read(F_FRAC)[c]
    assert a[0] - 0.5 == pytest.approx(change, rel=1e-9)


# --------------------------------------------------------------------------- IMM-CT
def immct_cfg(vis_max=0.01, dt=0.001):
    cfg = scenes.immct_config(radius=R, lb=(0, 0, 0), size=(1, 1, 1), dt=dt, three_phase=False, solver="IMMCT")
    req = cfg["IMMCT"]["Required"]
    req.update({"phaseRestDensity": [1000.0, 1000.0], "solventViscosity": 0.0, "diffusionCoefficientCf": 0.0,
                "divFreeThreshold": 10.0, "incompThreshold": 10.0, "solutionMaxViscosity": vis_max})
    return _no_gravity(cfg, "IMMCT")


def test_immct_velocity_gradient_of_linear_field_and_cd_closed_form():
    """v = A (x - x_c) with a traceless, non-symmetric A (no compression from
    the flow; the block's own compression only moves its faces), fractions
    (0.3, 0.7).  At the centre of a 15^3 block: L = sum_j V_j (v_j - v_i)
    outer gradW_ij = -V0 S_f A, and Cd = Cd0 + (1 - Cd0) a_1 V0 sum_{o != 0} W."""
    pos, idx = box_lattice(15, 15, 15, S_C)
    n = pos.shape[0]
    c = at(idx, 7, 7, 7)
    A = np.array([[0.0, 0.3, 0.0], [-0.1, 0.0, 0.2], [0.05, 0.0, 0.0]])
    vel = ((pos.astype(np.float64) - pos[c].astype(np.float64)) @ A.T).astype(np.float32)
    frac = np.tile(np.array([0.3, 0.7], np.float32), (n, 1))
    # the test's field is the FP32 one the oracle receives: use it in the expectation
    Aeff = A   # v is linear up to FP32 rounding of each velocity (relative 6e-8)
    o, p = run("IMMCT", immct_cfg(), pos, vel, frac)
    L = o.read(F_L)[c].reshape(3, 3)
    expect = -V0 * s_f(S_C) * Aeff
    np.testing.assert_allclose(L, expect, rtol=0, atol=2e-7 * np.abs(expect).max())
    assert np.abs(L - L.T).max() > 0.1 * np.abs(expect).max()     # not symmetrised or transposed
    a1 = float(np.float32(0.7))
    cd0 = float(np.float32(p.cd))
    blocking = a1 * V0 * sum(w(S_C * math.sqrt(a * a + b * b + cc * cc)) for a, b, cc in offsets(S_C)
                             if (a, b, cc) != (0, 0, 0))
    assert o.read(pyoracle.PROBE_BLOCKING)[c] == pytest.approx(blocking, rel=1e-12)
    assert o.read(F_CD)[c] == pytest.approx(cd0 + (1 - cd0) * blocking, rel=1e-12)


def test_immct_rheological_factors_at_special_fractions():
    """beta = e0 max(0, th - a)/(th + 1e-6) and kappa = (1 - a)^-1.2 of the
    polymer fraction a: beta(0) = e0 th/(th + 1e-6), beta(a >= th) = 0,
    kappa(0) = 1, kappa(1/2) = 2^1.2.  At rest (L = 0) the viscosity is
    base + (min(base kappa, max) - base) min(1, a/(th + 1e-6)): base at a = 0
    and base 2^1.2 at a = 1/2 with a large maximum."""
    s = 3.0 * 2.0 ** -6            # sparse lattice: cr < 1, nothing moves
    pos, idx = box_lattice(6, 6, 6, s)
    n = pos.shape[0]
    a_of = np.array([0.0, 0.4, 0.5, 0.25], np.float32)
    a = a_of[np.arange(n) % 4]
    frac = np.stack([1.0 - a, a], 1).astype(np.float32)
    o, p = run("IMMCT", immct_cfg(vis_max=100.0), pos, None, frac)
    e0, th = float(np.float32(p.thinning_exp0)), float(np.float32(p.rheology_threshold))
    base = 1000.0 * float(np.float32(p.solution_vis_base))
    beta, kap, vis = o.read(pyoracle.PROBE_BETA), o.read(pyoracle.PROBE_KAPPA_CT), o.read(F_VIS)
    np.testing.assert_allclose(o.read(F_L), 0.0, atol=1e-15)
    for i in range(4):
        k = i
        ai = float(a_of[i])
        if ai == 0.0:
            assert beta[k] == pytest.approx(e0 * th / (th + 1e-6), rel=1e-14)
            assert kap[k] == 1.0
            assert vis[k] == pytest.approx(base, rel=1e-14)
        if ai >= th:
            assert beta[k] == 0.0
        if ai == 0.5:
            assert kap[k] == pytest.approx(2.0 ** 1.2, rel=1e-14)
            assert vis[k] == pytest.approx(base * 2.0 ** 1.2, rel=1e-14)
        if ai == 0.25:
            assert beta[k] == pytest.approx(e0 * (th - 0.25) / (th + 1e-6), rel=1e-12)


def test_immct_polymer_stress_acceleration_at_slab_face():
    """A 9 x 9 x 4 slab at spacing 1.2 * 2r (cr < 1: no pressure), v = 0,
    uniform fractions (3/8, 5/8) and a uniform conformation Q != I, so every
    fluid particle carries the same stress S (read back and checked uniform).
    Interior sums of gradW vanish; a face particle gets the acceleration
    (2 a_p / rho) S g with g = sum_j gradW_ij over the layer below, equal on
    every phase (equal phase densities), so after the tick v = dt (2 a_p/rho) S g."""
    s = 3.0 * 2.0 ** -7
    nz = 4
    pos, idx = box_lattice(9, 9, nz, s)
    n = pos.shape[0]
    frac = np.tile(np.array([0.375, 0.625], np.float32), (n, 1))   # exact: rest density exactly 1000
    Q = np.array([[1.2, 0.1, 0.3], [0.1, 0.9, -0.2], [0.3, -0.2, 1.1]])

    def conformation(orc):
        orc.write(F_Q, np.tile(Q.ravel(), (n, 1)))

    o, p = run("IMMCT", immct_cfg(), pos, None, frac, before=conformation)
    dt = float(np.float32(p.time_step))
    rho = float(np.float32(p.phase_rest_density[0]))
    S = o.read(F_STRESS)
    fl = np.ones(n, bool)
    np.testing.assert_allclose(S, np.tile(S[0], (n, 1)), rtol=0, atol=1e-12 * np.abs(S[0]).max())
    S0 = S[0].reshape(3, 3)
    assert np.abs(S0).max() > 1e-3 and fl.all()
    g = sum(grad(-np.array(v, np.float64) * s) for v in offsets(s) if v[2] == -1)
    assert g[2] < 0 and abs(g[0]) < 1e-12 and abs(g[1]) < 1e-12   # W falls off: points into the slab
    a_p = 0.625
    expect = dt * (2.0 * a_p / rho) * (S0 @ g)
    top = at(idx, 4, 4, nz - 1)
    got = o.read(F_VEL)[top]
    np.testing.assert_allclose(got, expect, rtol=1e-10, atol=1e-12 * np.abs(expect).max())
    mid = at(idx, 4, 4, 1)   # a full stencil below and above: no net stress force
    assert np.abs(o.read(F_VEL)[mid]).max() <= 1e-12 * np.abs(expect).max()
