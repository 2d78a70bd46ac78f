"""SPH kernel functions of the oracle against closed forms
(KernelFunctions.cuh:13-69): values, continuity at q = 1/2, unit integral,
gradient = derivative of W, cohesion spline values."""
import math

import numpy as np
import pytest

import pyoracle

PI = 3.14159265


@pytest.mark.parametrize("h", [0.04, 0.02, 0.3])
def test_w_closed_forms(h):
    sigma = 8.0 / PI / h ** 3
    w0, g0, _ = pyoracle.kernel(h, [0, 0, 0])
    assert math.isclose(w0, sigma, rel_tol=1e-14)
    assert np.all(g0 == 0)      # q < 1e-6 -> zero gradient
    w_half, _, _ = pyoracle.kernel(h, [0.5 * h, 0, 0])
    assert math.isclose(w_half, sigma / 4, rel_tol=1e-12)
    w_out, g_out, _ = pyoracle.kernel(h, [0, 1.0001 * h, 0])
    assert w_out == 0 and np.all(g_out == 0)
    for q in (0.1, 0.3, 0.7, 0.9):
        w, _, _ = pyoracle.kernel(h, [0, 0, q * h])
        ref = sigma * (6 * q ** 3 - 6 * q ** 2 + 1) if q <= 0.5 else 2 * sigma * (1 - q) ** 3
        assert math.isclose(w, ref, rel_tol=1e-12)


def test_w_integrates_to_one():
    h = 0.04
    r = np.linspace(0, h, 20001)
    w = np.array([pyoracle.kernel(h, [x, 0, 0])[0] for x in r])
    integral = np.trapezoid(4 * np.pi * r ** 2 * w, r)
    # sigma uses the reference's pi literal, so the exact integral is pi / 3.14159265
    assert abs(integral - math.pi / PI) < 1e-7


def test_gradient_is_derivative_of_w():
    h = 0.04
    rng = np.random.default_rng(1)
    for _ in range(50):
        r = rng.uniform(-h, h, 3)
        if not 0.02 < np.linalg.norm(r) / h < 0.98 or abs(np.linalg.norm(r) / h - 0.5) < 1e-3:
            continue
        _, g, _ = pyoracle.kernel(h, r)
        eps = 1e-7
        fd = np.zeros(3)
        for a in range(3):
            e = np.zeros(3)
            e[a] = eps
            fd[a] = (pyoracle.kernel(h, r + e)[0] - pyoracle.kernel(h, r - e)[0]) / (2 * eps)
        np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9 * np.abs(fd).max())


def test_gradient_points_along_r_and_is_odd():
    h = 0.04
    r = np.array([0.01, -0.007, 0.013])
    _, g, _ = pyoracle.kernel(h, r)
    _, gm, _ = pyoracle.kernel(h, -r)
    np.testing.assert_allclose(g, -gm, rtol=1e-15)
    assert abs(np.dot(g, r) / (np.linalg.norm(g) * np.linalg.norm(r)) + 1.0) < 1e-12  # W decreasing


def test_cohesion_spline():
    h = 0.04
    c = 32.0 / PI / h ** 9
    for r in (0.3 * h, 0.5 * h, 0.51 * h, 0.9 * h):
        _, _, C = pyoracle.kernel(h, [r, 0, 0])
        if 2 * r > h:
            ref = c * (h - r) ** 3 * r ** 3
        else:
            ref = c * (2 * (h - r) ** 3 * r ** 3 - h ** 6 / 64)
        assert math.isclose(C, ref, rel_tol=1e-12)
    # continuous at r = h/2 and zero outside (0, h]
    a = pyoracle.kernel(h, [0.5 * h * (1 - 1e-9), 0, 0])[2]
    b = pyoracle.kernel(h, [0.5 * h * (1 + 1e-9), 0, 0])[2]
    assert math.isclose(a, b, rel_tol=1e-6)
    assert pyoracle.kernel(h, [1.01 * h, 0, 0])[2] == 0
    assert pyoracle.kernel(h, [0, 0, 0])[2] == 0
