"""Pins the oracle's generic focused-transport path (include/FocusSolver.hpp,
include/FocusCoefficient.hpp, include/LegencyModel.hpp) with closed forms:
central differences of trigonometric fields for grad ln|B|, the curvature
(b.grad)b, the two drift bases, b.grad ln|B|, bb:grad V and b.dV/dt; the kappa
fields and div(kappa_perp (I - bb)) against the generic Parker basis (pinned by
its own closed forms); in a uniform field the terms, the six-candidate local
time step and one step with injected noise against an independent
restatement of the Picard update over the oracle's slab-QLT D_mumu. No GPU is
needed."""
import numpy as np
import pytest

import kpt
import kpt_oracle as O
from helpers import uniform_edges

NX, NY = 32, 16
X0, X1, Y0, Y1 = -1.0, 1.0, -0.5, 0.5
KX = 2 * np.pi / (X1 - X0)
KY = 2 * np.pi / (Y1 - Y0)
HX, HY = (X1 - X0) / NX, (Y1 - Y0) / NY
SX, SY = np.sin(KX * HX) / HX, np.sin(KY * HY) / HY  # central difference of sin(k .)


def _grid():
    xe, ye = uniform_edges(NX, X0, X1), uniform_edges(NY, Y0, Y1)
    X, Y = np.meshgrid(0.5 * (xe[:-1] + xe[1:]), 0.5 * (ye[:-1] + ye[1:]))
    return xe, ye, X, Y


def _params(**kw):
    p = O.generic_focus_default_params()
    for k, v in kw.items():
        target = p.model if hasattr(p.model, k) and not hasattr(p, k) else p
        setattr(target, k, v)
    return p


def _field(planes, dvdt=None, **kw):
    xe, ye, _, _ = _grid()
    return O.GenericFocusField(xe, ye, planes, _params(**kw), dvdt=dvdt)


def _coef(f):
    return f.coefficients()[1:-1, 1:-1, :]


def test_defaults_match_reference(oracle_mod):
    p = O.generic_focus_default_params()
    assert p.spectral_index == 5.0 / 3.0 and p.resonance_regularization_h0 == 1e-6
    assert p.maximum_pitch_angle_cosine == 1.0 and p.minimum_magnetic_field_magnitude == 0.0
    assert p.model.courant == 0.3 and p.model.parallel_prefactor == 1.622


def test_direction_curvature_and_drift_bases(oracle_mod):
    _, _, X, Y = _grid()
    a = 0.3
    A = np.exp(a * np.sin(KX * X))          # |B| = A: grad ln|B| = (a Kx cos(Kx x), 0, 0) up to the stencil
    bx, by = np.cos(KY * Y), np.sin(KY * Y)  # b rotates with y
    vx = np.full(X.shape, 0.2)
    f = _field([A * bx, A * by, 0 * X, vx, 0 * X, 0 * X])
    c = _coef(f)
    np.testing.assert_allclose(c[..., 3], bx, atol=1e-14)
    np.testing.assert_allclose(c[..., 4], by, atol=1e-14)
    np.testing.assert_allclose(c[..., 18], A, rtol=1e-14)
    g = a * np.cos(KX * X) * SX
    # (b.grad)b = b_y d_y b: (-sin^2, sin cos) * S_y
    cx, cy = -by * by * SY, by * bx * SY
    # drift bases b x grad ln|B| / |B| and b x (b.grad)b / |B|
    np.testing.assert_allclose(c[..., 9:11], 0.0, atol=1e-13)
    np.testing.assert_allclose(c[..., 11], -by * g / A, atol=1e-12)
    np.testing.assert_allclose(c[..., 12:14], 0.0, atol=1e-13)
    np.testing.assert_allclose(c[..., 14], (bx * cy - by * cx) / A, atol=1e-12)
    np.testing.assert_allclose(c[..., 14], by * SY / A, atol=1e-12)
    np.testing.assert_allclose(c[..., 22], bx * g, atol=1e-12)           # b.grad ln|B|
    np.testing.assert_allclose(c[..., 21], 0.0, atol=1e-13)             # uniform V
    np.testing.assert_allclose(c[..., 23:25], 0.0, atol=1e-13)
    # ghosts are periodic images
    full = f.coefficients()
    np.testing.assert_array_equal(full[0, 1:-1], full[-2, 1:-1])
    np.testing.assert_array_equal(full[1:-1, 0], full[1:-1, -2])


def test_solar_wind_contractions(oracle_mod):
    _, _, X, Y = _grid()
    bx, by = np.cos(KY * Y), np.sin(KY * Y)
    v0, w0 = 0.4, 0.25
    vx, vy = v0 * np.sin(KX * X), w0 * np.cos(KY * Y)
    dxvx, dyvy = v0 * np.cos(KX * X) * SX, -w0 * np.sin(KY * Y) * SY
    t0, t1 = 0.7 * np.cos(KX * X), -0.3 * np.ones(X.shape)
    f = _field([bx, by, 0 * X, vx, vy, 0 * X], dvdt=[t0, t1, 0 * X])
    c = _coef(f)
    np.testing.assert_allclose(c[..., 21], dxvx + dyvy, atol=1e-12)
    np.testing.assert_allclose(c[..., 23], bx * bx * dxvx + by * by * dyvy, atol=1e-12)
    np.testing.assert_allclose(c[..., 24], bx * t0 + by * t1 + bx * vx * dxvx + by * vy * dyvy, atol=1e-12)
    # without dV/dt the partial-time projection drops out
    c0 = _coef(_field([bx, by, 0 * X, vx, vy, 0 * X]))
    np.testing.assert_allclose(c0[..., 24], bx * vx * dxvx + by * vy * dyvy, atol=1e-12)


@pytest.mark.parametrize("model", [0, 1])
def test_kappa_and_perpendicular_advection_match_parker_basis(oracle_mod, model):
    xe, ye, X, Y = _grid()
    planes = [np.cos(KY * Y), np.sin(KY * Y), 0.3 + 0.1 * np.sin(KX * X), 0.2 + 0.05 * np.cos(KX * X),
              0.1 * np.sin(KY * Y), 0 * X]
    kw = dict(perpendicular_model=model, perpendicular_parallel_ratio=0.1 if model else 0.0, sigma2=0.3,
              l2d=0.05, lslab=0.07)
    foc = _coef(_field(planes, **kw))
    pp = O.generic_default_params()
    for k, v in kw.items():
        setattr(pp, k, v)
    par = O.GenericField(xe, ye, planes, pp).coefficients()[1:-1, 1:-1, :]
    np.testing.assert_array_equal(foc[..., 19], par[..., 15])
    np.testing.assert_array_equal(foc[..., 20], par[..., 16])
    if model == 0:  # README model: the Parker secondary basis is div(kappa_perp (I - bb))
        np.testing.assert_allclose(foc[..., 15:18], par[..., 12:15], rtol=1e-13, atol=1e-15)
    else:
        # constant ratio r: div(r kappa_par (I - bb)) = div(kappa_par (r I + (1 - r) bb)) - div(kappa_par bb),
        # the Parker ratio basis minus the README model's parallel basis
        pp.perpendicular_model, pp.perpendicular_parallel_ratio = 0, 0.0
        readme = O.GenericField(xe, ye, planes, pp).coefficients()[1:-1, 1:-1, :]
        np.testing.assert_allclose(foc[..., 15:18], par[..., 9:12] - readme[..., 9:12], rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(foc[..., 25], 0.3)
    np.testing.assert_array_equal(foc[..., 26], 0.07)  # lc <= 0 -> lslab


def test_abort_cases(oracle_mod):
    _, _, X, Y = _grid()
    B = [np.ones(X.shape), 0 * X, 0 * X, np.full(X.shape, 0.1), 0 * X, 0 * X]
    _field(B)
    Bz = [b.copy() for b in B]
    Bz[0][3, 4] = 0.0
    with pytest.raises(ValueError):
        _field(Bz)                                    # direction undefined at |B| = 0
    with pytest.raises(ValueError):
        _field(B, minimum_magnetic_field_magnitude=1.0)  # |B| <= floor aborts the direction field


def _uniform(b0=1.3, v0=0.05, **kw):
    _, _, X, _ = _grid()
    one = np.ones(X.shape)
    return _field([b0 * one, 0 * X, 0 * X, v0 * one, 0 * X, 0 * X], **kw), b0, v0


@pytest.mark.parametrize("model", [0, 1])
def test_uniform_field_terms_and_local_dt(oracle_mod, model):
    f, b0, v0 = _uniform(perpendicular_model=model, perpendicular_parallel_ratio=0.2 if model else 0.0,
                         sigma2=0.2, lc=0.03)
    c = f.coefficients()[5, 7]
    kpar1, kperp1 = c[19], c[20]
    for p, mu in ((0.5, 0.3), (3.0, -0.8), (0.1, 0.0), (2.0, 1.0)):
        t = f.terms(0.1, -0.2, p, mu)
        g = np.sqrt(1 + p * p)
        v = p / g
        fperp = g ** (1 / 3) if model else g ** (1 / 9)
        assert t["kpar"] == pytest.approx(kpar1 * g ** (1 / 3), rel=1e-14)
        assert t["kperp"] == pytest.approx(kperp1 * fperp, rel=1e-14)
        assert t["u_x"] == pytest.approx(v * mu + v0, rel=1e-14, abs=1e-16)
        assert t["u_y"] == 0.0 and t["p_rate"] == 0.0
        assert t["keff_x"] == 0.0 and t["keff_y"] == pytest.approx(t["kperp"], rel=1e-15)
        D, dD = O.qlt_scattering(mu, b0 / g, v, 0.2, 0.03)
        assert t["D"] == pytest.approx(D, rel=1e-14, abs=1e-300)
        assert t["dD"] == pytest.approx(dD, rel=1e-14, abs=1e-300) and t["mu_rate"] == t["dD"]
        cands = [(0.5 * HY) ** 2 / (2 * t["kperp"]), HX / abs(v * mu + v0)]
        if dD != 0:
            cands.append(10 * 0.1 / abs(dD))
        if D > 0:
            cands.append(10 * 0.1 * 0.1 / (2 * D))
        if D > 0 and dD != 0:
            cands.append(10 * 2 * D / dD ** 2)
        assert t["dt_local"] == pytest.approx(min(cands), rel=1e-13)


def test_uniform_field_step_with_injected_noise(oracle_mod):
    f, b0, v0 = _uniform(sigma2=0.2, lc=0.03)
    n = 64
    rng = np.random.default_rng(5)
    parts = kpt.empty_particles(n)
    parts["x"][:] = rng.uniform(X0, X1, n)
    parts["y"][:] = rng.uniform(Y0, Y1, n)
    parts["p"][:] = np.exp(rng.uniform(np.log(0.2), np.log(5.0), n))
    parts["mu"][:] = rng.uniform(-1, 1, n)
    parts["id"][:] = np.arange(n)
    start = {k: v.copy() for k, v in parts.items()}
    noise = rng.standard_normal((n, 1, 4))
    dt_want = 0.3 * min(f.terms(parts["x"][i], parts["y"][i], parts["p"][i], parts["mu"][i])["dt_local"]
                        for i in range(n))
    assert f.time_step(parts) == pytest.approx(dt_want, rel=1e-15)
    dt, nf = f.step(parts, 0, noise=noise)
    assert nf == 0 and dt == pytest.approx(dt_want, rel=1e-15)
    for i in range(n):
        p, mu = start["p"][i], start["mu"][i]
        t = f.terms(start["x"][i], start["y"][i], p, mu)
        g = np.sqrt(1 + p * p)
        v = p / g
        # b = x-hat: the perpendicular increment has no x part
        want_x = start["x"][i] + dt * (v * mu + v0)
        want_x = (want_x - X0) % (X1 - X0) + X0
        want_y = start["y"][i] + np.sqrt(2 * t["kperp"] * dt) * noise[i, 0, 1]
        want_y = (want_y - Y0) % (Y1 - Y0) + Y0
        assert parts["x"][i] == pytest.approx(want_x, abs=1e-13)
        assert parts["y"][i] == pytest.approx(want_y, abs=1e-13)
        assert parts["p"][i] == p
        # independent three-pass Picard update over the oracle's D_mumu
        trial = mu0 = O.reflect_mu(mu)
        for _ in range(3):
            m = O.reflect_mu(trial)
            D, dD = O.qlt_scattering(m, b0 / g, v, 0.2, 0.03)
            trial = O.reflect_mu(mu0 + dt * dD + (np.sqrt(2 * D * dt) if D > 0 else 0.0) * noise[i, 0, 3])
        assert parts["mu"][i] == pytest.approx(trial, abs=1e-13)


def test_momentum_and_focusing_rates_closed_form(oracle_mod):
    _, _, X, Y = _grid()
    v0 = 0.3
    vx = v0 * np.sin(KX * X)
    f = _field([np.ones(X.shape), 0 * X, 0 * X, vx, 0 * X, 0 * X], sigma2=0.2, lc=0.03)
    i, j = 5, 3
    x, y = X[j, i], Y[j, i]  # a cell centre: the interpolation returns the node
    dvx = v0 * np.cos(KX * x) * SX
    for p, mu in ((0.7, 0.4), (2.0, -0.6)):
        g = np.sqrt(1 + p * p)
        v = p / g
        t = f.terms(x, y, p, mu)
        bdv = v0 * np.sin(KX * x) * dvx
        assert t["p_rate"] == pytest.approx(-p * (mu * mu * dvx + mu * bdv / v), rel=1e-12)
        D, dD = O.qlt_scattering(mu, 1.0 / g, v, 0.2, 0.03)
        foc = 0.5 * (1 - mu * mu) * (mu * dvx - 3 * mu * dvx - 2 * bdv / v)
        assert t["mu_rate"] == pytest.approx(foc + dD, rel=1e-12)


def test_zero_momentum_fails(oracle_mod):
    f, _, _ = _uniform()
    assert f.terms(0.0, 0.0, 0.0, 0.5) is None  # the reference aborts at zero speed
    parts = kpt.empty_particles(2)
    parts["p"][:] = [0.0, 1.0]
    parts["mu"][:] = 0.5
    parts["id"][:] = [0, 1]
    dt, nf = f.step(parts, 0, seed=3)
    assert nf == 1 and parts["status"].tolist() == [2, 0] and dt > 0


@pytest.mark.parametrize("model", [0, 1])
def test_stretched_grid_sampling_and_local_dt(oracle_mod, model):
    """Both axes stretched (spacing ratios 6 and 10), B = b0 x-hat and V_x linear
    in the cell-centre y coordinate: between the first and last cell centres the
    sampled V_x is that line (linear interpolation between the face midpoints), so
    u_x = v mu + V_x(y) exactly, and the local time step takes the face-to-face
    widths of the particle's own cell."""
    import workloads as W
    xe, ye = W.stretched_edges(8, X0, X1, ratio=6.0), W.stretched_edges(NY, Y0, Y1, ratio=10.0)
    nx, ny = len(xe) - 1, len(ye) - 1
    yc = 0.5 * (ye[:-1] + ye[1:])
    one = np.ones((ny, nx))
    b0 = 1.3
    vx = (0.05 + 0.02 * yc)[:, None] * one
    p = _params(perpendicular_model=model, perpendicular_parallel_ratio=0.2 if model else 0.0, sigma2=0.2,
                lc=0.03)
    f = O.GenericFocusField(xe, ye, [b0 * one, 0 * one, 0 * one, vx, 0 * one, 0 * one], p)
    rng = np.random.default_rng(9)
    widths = set()
    for x, y, pm, mu in zip(rng.uniform(X0, X1, 30), rng.uniform(yc[0], yc[-1], 30), np.exp(rng.uniform(-2, 1, 30)),
                            rng.uniform(-0.95, 0.95, 30)):
        t = f.terms(x, y, pm, mu)
        g = np.sqrt(1 + pm * pm)
        v = pm / g
        ux = v * mu + 0.05 + 0.02 * y
        assert t["u_x"] == pytest.approx(ux, rel=1e-13, abs=1e-16)
        i = np.searchsorted(xe, x, side="right") - 1
        j = np.searchsorted(ye, y, side="right") - 1
        hx, hy = xe[i + 1] - xe[i], ye[j + 1] - ye[j]
        D, dD = O.qlt_scattering(mu, b0 / g, v, 0.2, 0.03)
        cands = [(0.5 * hy) ** 2 / (2 * t["kperp"]), hx / abs(ux)]
        if dD != 0:
            cands.append(10 * 0.1 / abs(dD))
        if D > 0:
            cands.append(10 * 0.1 * 0.1 / (2 * D))
        if D > 0 and dD != 0:
            cands.append(10 * 2 * D / dD ** 2)
        assert t["dt_local"] == pytest.approx(min(cands), rel=1e-13), (x, y)
        widths.add(round(hy, 12))
    assert len(widths) > 6
