"""Pins the oracle's generic Parker path (include/ParkerSolver.hpp,
include/ParkerCoefficient.hpp, include/LegencyModel.hpp, include/FieldOperator.hpp)
with closed forms: central differences of trigonometric fields for curl(B/B^2)
and div V, the kappa scaling laws and the README constants at unit inputs, the
projection structure of the div(kappa_ij) bases, and the global time step and
the roles of the four Gaussians in one step. No GPU is needed."""
import numpy as np
import pytest

import kpt
import kpt_oracle as O
import workloads as W
from helpers import uniform_edges

NX, NY = 32, 16
X0, X1, Y0, Y1 = -1.0, 1.0, -0.5, 0.5
K = 2 * np.pi / (X1 - X0)   # one period across x
L = 2 * np.pi / (Y1 - Y0)


def _grid():
    xe, ye = uniform_edges(NX, X0, X1), uniform_edges(NY, Y0, Y1)
    xc = 0.5 * (xe[:-1] + xe[1:])
    yc = 0.5 * (ye[:-1] + ye[1:])
    X, Y = np.meshgrid(xc, yc)
    return xe, ye, X, Y


def _field(planes, **kw):
    xe, ye, _, _ = _grid()
    p = O.generic_default_params()
    for k, v in kw.items():
        setattr(p, k, v)
    return O.GenericField(xe, ye, planes, p)


def _coef(f):
    return f.coefficients()[1:-1, 1:-1, :]  # physical nodes [ny][nx][18]


def _const(shape, v):
    return np.full(shape, float(v))


def test_curl_of_b_over_b2_closed_forms(oracle_mod):
    xe, ye, X, Y = _grid()
    h = (X1 - X0) / NX
    damp = np.sin(K * h) / (K * h)  # central difference of sin(Kx)
    # B = Bz(x) z-hat with Bz/B^2 = 2 + sin(Kx): curl_y = -d/dx(2 + sin Kx)
    bz = 1.0 / (2.0 + np.sin(K * X))
    f = _field([0 * X, 0 * X, bz, _const(X.shape, 0.3), 0 * X, 0 * X])
    c = _coef(f)
    np.testing.assert_allclose(c[..., 7], -K * np.cos(K * X) * damp, atol=1e-12)
    np.testing.assert_allclose(c[..., 6], 0.0, atol=1e-12)
    np.testing.assert_allclose(c[..., 8], 0.0, atol=1e-12)
    # |B| = 1, B = (cos Kx, sin Kx, 0): curl_z = d/dx sin(Kx)
    f = _field([np.cos(K * X), np.sin(K * X), 0 * X, _const(X.shape, 0.3), 0 * X, 0 * X])
    c = _coef(f)
    np.testing.assert_allclose(c[..., 8], K * np.cos(K * X) * damp, atol=1e-12)
    np.testing.assert_allclose(c[..., 6:8], 0.0, atol=1e-12)


def test_divergence_of_v_closed_form(oracle_mod):
    _, _, X, Y = _grid()
    hx, hy = (X1 - X0) / NX, (Y1 - Y0) / NY
    vx = 1.0 + 0.2 * np.sin(K * X)
    vy = 0.3 * np.cos(L * Y)
    f = _field([_const(X.shape, 1.0), 0 * X, 0 * X, vx, vy, 0 * X])
    want = 0.2 * K * np.cos(K * X) * np.sin(K * hx) / (K * hx) - 0.3 * L * np.sin(L * Y) * np.sin(L * hy) / (L * hy)
    np.testing.assert_allclose(_coef(f)[..., 17], want, atol=1e-12)


def test_kappa_unit_values_and_scaling(oracle_mod):
    shape = (NY, NX)

    def kappas(b=1.0, v=1.0, **kw):
        f = _field([_const(shape, b), 0 * _const(shape, 0), 0 * _const(shape, 0), _const(shape, v),
                    0 * _const(shape, 0), 0 * _const(shape, 0)], **kw)
        c = _coef(f)
        return c[0, 0, 15], c[0, 0, 16]

    kpar, kperp = kappas(sigma2=1.0, lslab=1.0, l2d=1.0)
    # README constants at unit V, Lc, Omega, sigma2: kpar = 1.622,
    # kperp = (1/3) (3 kpar)^(1/3) (0.198)^(2/3)
    assert kpar == pytest.approx(1.622, rel=1e-14)
    assert kperp == pytest.approx((3 * 1.622) ** (1 / 3) * 0.198 ** (2 / 3) / 3, rel=1e-14)
    base = kappas()[0]
    assert kappas(v=2.0)[0] == pytest.approx(base * 2 ** (4 / 3), rel=1e-13)
    assert kappas(b=8.0)[0] == pytest.approx(base / 2, rel=1e-13)
    assert kappas(sigma2=1.0)[0] == pytest.approx(base * 0.5, rel=1e-13)
    assert kappas(lslab=8.0)[0] == pytest.approx(base * 4, rel=1e-13)
    assert kappas(lc=8.0)[0] == pytest.approx(base * 4, rel=1e-13)  # explicit Lc wins over lslab
    assert kappas(charge=-8.0)[0] == pytest.approx(base / 2, rel=1e-13)  # |q| in Omega
    # README kperp ~ V^(2/3) kpar^(1/3) ~ V^(10/9); constant ratio: kperp = r kpar
    assert kappas(v=2.0)[1] == pytest.approx(kappas()[1] * 2 ** (10 / 9), rel=1e-13)
    kp, kq = kappas(perpendicular_model=1, perpendicular_parallel_ratio=0.05)
    assert kq == pytest.approx(0.05 * kp, rel=1e-15)


def test_divergence_bases_project_onto_b(oracle_mod):
    """kappa_par1 varies with x through V_x(x) (kappa ~ V^(4/3)). With b along x
    the README parallel basis is d_x kappa_par1 and the perpendicular one is 0;
    with b along y the roles swap (perpendicular basis = d_x kappa_perp1). The
    constant-ratio basis is d_x kappa_par1 (b = x) and r d_x kappa_par1 (b = y)."""
    _, _, X, _ = _grid()
    vx = 1.0 + 0.3 * np.sin(K * X)
    zero = 0 * X
    h = (X1 - X0) / NX

    def ddx(a):
        return (np.roll(a, -1, axis=1) - np.roll(a, 1, axis=1)) / (2 * h)

    for bvec in ((1.0, 0.0), (0.0, 1.0)):
        f = _field([_const(X.shape, bvec[0]), _const(X.shape, bvec[1]), zero, vx, zero, zero])
        c = _coef(f)
        dpar, dperp = ddx(c[..., 15]), ddx(c[..., 16])
        if bvec[0] == 1.0:
            np.testing.assert_allclose(c[..., 9], dpar, rtol=1e-10, atol=1e-12)
            np.testing.assert_allclose(c[..., 12:15], 0.0, atol=1e-12)
        else:
            np.testing.assert_allclose(c[..., 9:12], 0.0, atol=1e-12)
            np.testing.assert_allclose(c[..., 12], dperp, rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(c[..., 10:12] if bvec[0] == 1.0 else c[..., 13:15], 0.0, atol=1e-12)
        r = 0.07
        g = _field([_const(X.shape, bvec[0]), _const(X.shape, bvec[1]), zero, vx, zero, zero],
                   perpendicular_model=1, perpendicular_parallel_ratio=r)
        cg = _coef(g)
        want = ddx(cg[..., 15]) * (1.0 if bvec[0] == 1.0 else r)
        np.testing.assert_allclose(cg[..., 9], want, rtol=1e-10, atol=1e-12)
        np.testing.assert_array_equal(cg[..., 9:12], cg[..., 12:15])  # secondary = primary


def test_abort_cases_refuse_the_field(oracle_mod):
    _, _, X, _ = _grid()
    zero = 0 * X
    with pytest.raises(ValueError):  # |B| = 0 somewhere, no floor
        _field([np.where(X > 0, 1.0, 0.0), zero, zero, _const(X.shape, 0.3), zero, zero])
    with pytest.raises(ValueError):  # V = 0 somewhere, no floor
        _field([_const(X.shape, 1.0), zero, zero, np.where(X > 0, 0.3, 0.0), zero, zero])
    _field([_const(X.shape, 1.0), zero, zero, np.where(X > 0, 0.3, 0.0), zero, zero],
           minimum_solar_wind_speed=0.01)  # the floor admits it


def _uniform_case(v0=0.3, ratio=0.05, p=1.0):
    _, _, X, _ = _grid()
    zero = 0 * X
    f = _field([_const(X.shape, 1.0), zero, zero, _const(X.shape, v0), zero, zero],
               perpendicular_model=1, perpendicular_parallel_ratio=ratio)
    parts = kpt.empty_particles(3)
    parts["x"][:] = [0.1, -0.7, 0.95]
    parts["y"][:] = [0.0, 0.2, -0.45]
    parts["p"][:] = p
    parts["id"][:] = [0, 1, 2]
    return f, parts


def test_global_time_step_closed_form(oracle_mod):
    """B = x-hat, V = v0 x-hat uniform, constant ratio r: u = (v0, 0, 0),
    kappa_par = kpar1 gamma^(1/3), kappa_perp = r kappa_par, div V = 0, so
    dt = 0.3 min((dx/2)^2 / (2 kappa_par), (dy/2)^2 / (2 kappa_perp), dx / v0)."""
    f, parts = _uniform_case(p=2.0)
    kpar1 = _coef(f)[0, 0, 15]
    gamma = np.sqrt(1 + 2.0 ** 2)
    kpar = kpar1 * gamma ** (1 / 3)
    dx, dy = (X1 - X0) / NX, (Y1 - Y0) / NY
    want = 0.3 * min((0.5 * dx) ** 2 / (2 * kpar), (0.5 * dy) ** 2 / (2 * 0.05 * kpar), dx / 0.3)
    assert f.time_step(parts) == pytest.approx(want, rel=1e-13)
    parts["status"][:] = 2
    assert f.time_step(parts) == 0.0  # no alive particle


def test_step_noise_roles_closed_form(oracle_mod):
    """Same uniform case with injected Gaussians (g_x, g_y, g_z, g_par): with b = x
    the isotropic x part is replaced by the field-aligned draw, so
    dx = dt v0 + sqrt(2 kappa_par dt) g_par, dy = sqrt(2 kappa_perp dt) g_y, p fixed."""
    f, parts = _uniform_case(p=1.0)
    start = {k: v.copy() for k, v in parts.items()}
    noise = np.array([[[0.4, -1.1, 0.8, 0.25]], [[-0.3, 0.5, 2.0, -1.5]], [[1.2, 0.0, -0.7, 0.9]]])
    dt, fails = f.step(parts, 0, noise=noise)
    assert fails == 0 and dt > 0
    kpar = _coef(f)[0, 0, 15] * np.sqrt(2.0) ** (1 / 3)
    kperp = 0.05 * kpar
    for i in range(3):
        x = start["x"][i] + dt * 0.3 + np.sqrt(2 * kpar * dt) * noise[i, 0, 3]
        y = start["y"][i] + np.sqrt(2 * kperp * dt) * noise[i, 0, 1]
        x = (x - X0) % (X1 - X0) + X0
        y = (y - Y0) % (Y1 - Y0) + Y0
        assert parts["x"][i] == pytest.approx(x, abs=1e-14)
        assert parts["y"][i] == pytest.approx(y, abs=1e-14)
    np.testing.assert_array_equal(parts["p"], start["p"])


def test_step_momentum_and_philox_normals(oracle_mod):
    """p_new = p - dt p div V / 3 at the sampled div V; Philox/Box-Muller normals
    have zero mean and unit variance."""
    _, _, X, Y = _grid()
    zero = 0 * X
    vx = 0.3 + 0.1 * np.sin(K * X)
    f = _field([_const(X.shape, 1.0), zero, zero, vx, zero, zero], perpendicular_model=1,
               perpendicular_parallel_ratio=0.05)
    n = 20000
    rng = np.random.default_rng(1)
    parts = kpt.empty_particles(n)
    parts["x"][:] = rng.uniform(X0, X1, n)
    parts["y"][:] = rng.uniform(Y0, Y1, n)
    parts["p"][:] = 1.0
    parts["id"][:] = np.arange(n)
    start = {k: v.copy() for k, v in parts.items()}
    dt, _ = f.step(parts, 3, seed=5)
    divv = np.array([f.sample(start["x"][i], start["y"][i])[17] for i in range(50)])
    np.testing.assert_allclose(parts["p"][:50], 1.0 - dt * divv / 3.0, rtol=1e-14)
    # y increments are sqrt(2 kperp dt) g_y with b = x: on the uniform field
    # (kappa_perp the same everywhere) the standardised increments are N(0, 1)
    f, _ = _uniform_case()
    parts = {k: v.copy() for k, v in start.items()}
    dt, _ = f.step(parts, 3, seed=5)
    kperp = 0.05 * _coef(f)[0, 0, 15] * np.sqrt(2.0) ** (1 / 3)
    dy = parts["y"] - start["y"]
    dy = (dy + 0.5) % 1.0 - 0.5
    g = dy / np.sqrt(2 * kperp * dt)
    assert abs(np.mean(g)) < 4 / np.sqrt(n)
    assert abs(np.var(g) - 1.0) < 0.05


def _stretched_edges():
    """Faces refined toward the middle of each axis (spacing ratio 6 in x, 10 in y)."""
    return W.stretched_edges(8, X0, X1, ratio=6.0), W.stretched_edges(NY, Y0, Y1, ratio=10.0)


def test_stretched_time_step_uses_the_particles_face_width(oracle_mod):
    """Stretched grid, uniform B = x-hat and V = v0 x-hat: the terms do not depend
    on position, so dt = 0.3 min((hx/2)^2 / (2 kappa_par), (hy/2)^2 / (2 kappa_perp),
    hx / v0) with (hx, hy) the face-to-face widths of the particle's own cell
    (StoredCoordinateGrid::physical_cell_width), particle by particle."""
    xe, ye = _stretched_edges()
    nx, ny = len(xe) - 1, len(ye) - 1
    one = np.ones((ny, nx))
    p = O.generic_default_params()
    p.perpendicular_model, p.perpendicular_parallel_ratio = 1, 0.05
    f = O.GenericField(xe, ye, [one, 0 * one, 0 * one, 0.3 * one, 0 * one, 0 * one], p)
    kpar = f.coefficients()[2, 2, 15] * np.sqrt(1 + 2.0 ** 2) ** (1 / 3)
    rng = np.random.default_rng(5)
    widths, binding = set(), 0
    for x, y in zip(rng.uniform(X0, X1, 40), rng.uniform(Y0, Y1, 40)):
        parts = kpt.empty_particles(1)
        parts["x"][:], parts["y"][:], parts["p"][:] = x, y, 2.0
        i = np.searchsorted(xe, x, side="right") - 1
        j = np.searchsorted(ye, y, side="right") - 1
        hx, hy = xe[i + 1] - xe[i], ye[j + 1] - ye[j]
        cands = [(0.5 * hx) ** 2 / (2 * kpar), (0.5 * hy) ** 2 / (2 * 0.05 * kpar), hx / 0.3]
        assert f.time_step(parts) == pytest.approx(0.3 * min(cands), rel=1e-13), (x, y)
        widths.add(round(hy, 12))
        binding += int(np.argmin(cands) == 1)
    assert len(widths) >= 6 and binding > 10  # the y width binds in cells of many widths (pairs by symmetry)


def test_stretched_derivatives_at_cell_centres(oracle_mod):
    """Cell-centred values on a stretched grid sit at the face midpoints: for V_y
    linear in the cell-centre coordinate, div V (three-point formula at unequal
    spacing, exact for linear data) equals the slope in every interior row."""
    xe, ye = _stretched_edges()
    nx, ny = len(xe) - 1, len(ye) - 1
    yc = 0.5 * (ye[:-1] + ye[1:])
    one = np.ones((ny, nx))
    vy = (0.2 + 0.7 * yc)[:, None] * one
    p = O.generic_default_params()
    p.minimum_solar_wind_speed = 1e-3
    f = O.GenericField(xe, ye, [one, 0 * one, 0.5 * one, 0.3 * one, vy, 0 * one], p)
    div = f.coefficients()[1:-1, 1:-1, 17]
    np.testing.assert_allclose(div[1:-1], 0.7, rtol=1e-12)
    assert np.abs(div[0] - 0.7).max() > 0.1  # the periodic wrap breaks linearity at the ends


def test_nearly_uniform_faces_take_the_stretched_path_to_the_same_result(oracle_mod):
    """Faces moved by 1e-9 (far beyond the uniform-spacing test's 1024 ulp) switch
    the oracle to its stored-face path; the basis, the time step and three
    injected-noise steps then agree with the uniform grid's to 1e-6."""
    _, _, X, Y = _grid()
    xe, ye = uniform_edges(NX, X0, X1), uniform_edges(NY, Y0, Y1)
    xs, ys = xe.copy(), ye.copy()
    xs[1:-1] += 1e-9 * np.sin(np.arange(1, NX))
    ys[1:-1] += 1e-9 * np.cos(np.arange(1, NY))
    planes = [1.0 + 0.2 * np.sin(K * X), 0.3 * np.cos(L * Y), 0 * X + 0.4, 0.3 + 0.1 * np.sin(K * X + L * Y),
              0.05 * np.cos(K * X), 0 * X]
    p = O.generic_default_params()
    p.minimum_solar_wind_speed = 1e-3
    a, b = O.GenericField(xe, ye, planes, p), O.GenericField(xs, ys, planes, p)
    ca, cb = a.coefficients(), b.coefficients()
    scale = np.abs(ca).max(axis=(0, 1)) + 1e-300
    assert (np.abs(ca - cb) / scale).max() < 1e-6
    rng = np.random.default_rng(2)
    pa = kpt.empty_particles(500)
    pa["x"][:], pa["y"][:] = rng.uniform(X0, X1, 500), rng.uniform(Y0, Y1, 500)
    pa["p"][:], pa["id"][:] = np.exp(rng.uniform(-1, 2, 500)), np.arange(500)
    pb = {k: v.copy() for k, v in pa.items()}
    assert b.time_step(pb) == pytest.approx(a.time_step(pa), rel=1e-6)
    noise = rng.standard_normal((500, 3, 4))
    for step in range(3):
        da, _ = a.step(pa, step, noise=noise)
        db, _ = b.step(pb, step, noise=noise)
        assert db == pytest.approx(da, rel=1e-6)
    for k, w in (("x", X1 - X0), ("y", Y1 - Y0)):
        d = np.abs(pa[k] - pb[k])
        assert (np.minimum(d, w - d) / w).max() < 1e-6, k
    np.testing.assert_allclose(pb["p"], pa["p"], rtol=1e-6)
