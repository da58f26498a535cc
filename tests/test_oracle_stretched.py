"""Pins the oracle's stretched-grid path (non-uniform faces, which the driver's
reader rejects; the build puts value k at face k): bilinear sampling reproduces
fields linear in y exactly anywhere between the faces; the packed y-derivatives
use the three-point formula, exact for quadratics at every interior face; the
settings drop the half-cell shift on a stretched axis; and a grid whose faces
differ from uniform by 1e-9 reproduces the uniform driver push (interpolation,
derivatives and the dt limiter's cell width) to that order. No GPU is needed."""
import numpy as np
import pytest

import kpt
import kpt_oracle as O
import workloads as W
from helpers import random_particles

NX, NY = 16, 24


def _grid(ratio=10.0):
    xe = np.linspace(-1.0, 1.0, NX + 1)
    ye = W.stretched_edges(NY, -0.5, 0.5, ratio)
    X, Y = np.meshgrid(xe[:-1], ye[:-1])  # values at the faces
    return xe, ye, X, Y


def test_stretched_edges_spacing():
    e = W.stretched_edges(512)
    d = np.diff(e)
    assert e[0] == -0.5 and e[-1] == 0.5 and np.all(d > 0)
    np.testing.assert_allclose(e, -e[::-1], atol=1e-15)
    assert 9.0 < d.max() / d.min() <= 10.0            # continuous ratio 10
    assert d.argmin() in (255, 256) and d.argmax() in (0, 511)


def test_settings_drop_the_shift_on_a_stretched_axis():
    xe, ye, _, _ = _grid()
    s = kpt.make_device_settings(kpt.default_host_settings(), xe, ye)
    assert s.interpolation_shift_x == pytest.approx(0.5 * 2.0 / NX) and s.interpolation_shift_y == 0.0
    assert s.y_lower == -0.5 and s.y_upper == 0.5


def test_linear_fields_sample_exactly(oracle_mod):
    xe, ye, X, Y = _grid()
    a, b = 0.3, 1.7
    one = np.ones_like(X)
    f = O.Field(xe, ye, [a + b * Y, 0.2 * one, 0.1 * one, 0.05 * one, -0.02 * one, 0 * one], kpt.MODEL_PARKER16)
    rng = np.random.default_rng(3)
    for qy in rng.uniform(ye[1], ye[-3], 40):          # away from the periodic seam
        qx = rng.uniform(-0.9, 0.9)
        c = f.sample(qx, qy)
        assert c[2] == pytest.approx(a + b * qy, rel=1e-13, abs=1e-15)   # Bx
        assert c[9] == pytest.approx(b, rel=1e-11)                        # dBx/dy


def test_three_point_derivative_exact_for_quadratics(oracle_mod):
    xe, ye, X, Y = _grid()
    one = np.ones_like(X)
    f = O.Field(xe, ye, [0.5 * one, Y * Y, 0.1 * one, 0.05 * one, 0.3 * Y * Y - Y, 0 * one],
                kpt.MODEL_PARKER16)
    c = f.coefficients()[1:-1, 1:-1, :]  # physical nodes [ny][nx][16]
    yk = ye[:-1]
    inner = slice(1, NY - 1)
    np.testing.assert_allclose(c[inner, :, 11], (2 * yk[inner])[:, None] * np.ones(NX), rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(c[inner, :, 7], (0.6 * yk[inner] - 1.0)[:, None] * np.ones(NX), rtol=1e-11,
                               atol=1e-13)
    np.testing.assert_allclose(c[:, :, 8], 0.0, atol=1e-14)                 # d Bx / dx
    # |B| = sqrt(0.25 + y^4 + 0.01): its derivative is not a quadratic; central
    # estimate within the three-point truncation error
    bm = np.sqrt(0.26 + yk ** 4)
    np.testing.assert_allclose(c[inner, 0, 15], (2 * yk ** 3 / bm)[inner], atol=2e-2)


@pytest.mark.parametrize("adaptive", [False, True])
def test_nearly_uniform_stretched_grid_reproduces_uniform_push(oracle_mod, adaptive):
    nx, ny = 32, 16
    xe, ye_u, planes = W.harris_field(nx, ny)
    ye_s = ye_u.copy()
    ye_s[1:-1] += 1e-9 * np.sin(np.arange(1, ny))  # breaks uniformity, moves faces by 1e-9
    h = W.parker_harris_host_settings()
    if adaptive:
        h.dt_min_rel, h.dt_max_rel = 1e-5, 1e-2
    s_u = kpt.make_device_settings(h, xe, ye_u)
    s_s = kpt.make_device_settings(h, xe, ye_s)
    assert s_s.interpolation_shift_y == 0.0 and s_u.interpolation_shift_y > 0.0
    rng = np.random.default_rng(5)
    parts = random_particles(512, s_u, rng, p_range=(0.5, 3.0))
    runs = []
    for ye, s in ((ye_u, s_u), (ye_s, s_s)):
        f = O.Field(xe, ye, planes, kpt.MODEL_PARKER16)
        p = {k: v.copy() for k, v in parts.items()}
        steps, fails = O.push_interval(f, f, s, 0.0, 0, 7, p)
        runs.append((steps, fails, p))
    (n_u, f_u, a), (n_s, f_s, b) = runs
    assert f_u == f_s == 0 and n_u == n_s
    for k in ("x", "y", "p", "t"):
        np.testing.assert_allclose(b[k], a[k], rtol=1e-6, atol=1e-7, err_msg=k)
    np.testing.assert_allclose(b["step_dt"], a["step_dt"], rtol=1e-5, err_msg="step_dt")
