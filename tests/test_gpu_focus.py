"""GPU parity of the generic focused-transport solver (include/FocusSolver.hpp)
against the oracle: the packed 27-component basis at every stored node (with
and without dV/dt), the global time step and ten injected-noise steps of x, y,
p and mu to 1e-12, on a uniform grid and on a grid stretched along both axes; with Philox, the perpendicular spread against 2 kappa_perp t;
zero-momentum particles marked and counted; the reference's abort cases
rejected at build time."""
import numpy as np
import pytest

import kpt
import kpt_oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def _case(model, n=4096, seed=3, stretched=False):
    if stretched:
        # both axes stretched: x by 4 toward x = 0, y by 10 toward the sheet
        ye = W.stretched_edges(32, ratio=10.0)
        _, ye, planes = W.island_field(64, 32, seed=seed, y_edges=ye)
        xe = W.stretched_edges(64, -1.0, 1.0, ratio=4.0)
    else:
        xe, ye, planes = W.island_field(64, 32, seed=seed)
    planes[2] = planes[2] + 0.2  # guide field: |B| > 0 everywhere
    p = kpt.generic_focus_default_params()
    p.model.perpendicular_model = model
    p.model.perpendicular_parallel_ratio = 0.05 if model == 1 else 0.0
    p.model.minimum_solar_wind_speed = 1e-3
    p.model.sigma2, p.model.lslab, p.model.l2d, p.model.lc = 0.1, 0.05, 0.02, 0.02
    p.maximum_pitch_angle_cosine = 0.995
    rng = np.random.default_rng(seed)
    parts = kpt.empty_particles(n)
    parts["x"][:] = rng.uniform(xe[0], xe[-1], n)
    parts["y"][:] = rng.uniform(ye[0], ye[-1], n)
    parts["p"][:] = np.exp(rng.uniform(np.log(0.5), np.log(50.0), n))
    parts["mu"][:] = rng.uniform(-1, 1, n)
    parts["id"][:] = np.arange(n)
    parts["status"][::17] = 2
    dvdt = [0.01 * np.cos(np.pi * planes[0]), 0.02 * planes[1], 0 * planes[2]]
    return xe, ye, planes, p, parts, dvdt


def _close(got, want, scale_axis=(0, 1)):
    scale = np.maximum(np.abs(want).max(axis=scale_axis), 1e-300)
    err = np.abs(got - want) / scale
    return err.max(), np.unravel_index(err.argmax(), err.shape)


@pytest.mark.parametrize("model,with_dvdt,stretched", [(0, False, False), (1, True, False), (0, True, True),
                                                        (1, False, True)])
def test_focus_basis_and_steps_match_oracle(model, with_dvdt, stretched):
    xe, ye, planes, p, parts, dvdt = _case(model, stretched=stretched)
    dv = dvdt if with_dvdt else None
    n = len(parts["x"])
    of = O.GenericFocusField(xe, ye, planes, p, dvdt=dv)
    noise = np.random.default_rng(7).standard_normal((n, 10, 4))
    with kpt.Context(n, seed=11) as c:
        c.field_upload(kpt.SLOT_LOWER, 16, xe, ye, planes)
        c.focus_build(p, kpt.SLOT_LOWER, dvdt=dv)
        err, where = _close(c.focus_coefficients(kpt.SLOT_LOWER, 64, 32), of.coefficients())
        assert err < 1e-12, (err, where)
        c.upload(parts)
        assert c.focus_time_step(p) == pytest.approx(of.time_step(parts), rel=1e-12)
        c.set_noise(noise)
        want = {k: v.copy() for k, v in parts.items()}
        for step in range(10):
            dt, nf = c.focus_step(p, step)
            dto, nfo = of.step(want, step, noise=noise)
            assert nf == nfo == 0
            assert dt == pytest.approx(dto, rel=1e-12)
        got = c.download()
    for k, w in (("x", xe[-1] - xe[0]), ("y", ye[-1] - ye[0])):
        d = np.abs(got[k] - want[k])
        d = np.minimum(d, w - d) / w
        assert d.max() < 1e-12, k
    np.testing.assert_allclose(got["p"], want["p"], rtol=1e-12)
    np.testing.assert_allclose(got["mu"], want["mu"], rtol=0, atol=1e-11)
    assert np.all(np.abs(got["mu"]) <= 0.995)
    np.testing.assert_array_equal(got["status"], want["status"])


def test_focus_philox_perpendicular_spread_and_failures():
    """Uniform B along x: y spreads with variance 2 kappa_perp t; a zero-momentum
    particle is marked (the reference aborts there) and counted once."""
    nx, ny = 32, 16
    xe, ye = np.linspace(-1, 1, nx + 1), np.linspace(-0.5, 0.5, ny + 1)
    one, zero = np.ones((ny, nx)), np.zeros((ny, nx))
    planes = [one, zero, zero, 0.05 * one, zero, zero]
    p = kpt.generic_focus_default_params()
    p.model.perpendicular_model, p.model.perpendicular_parallel_ratio = 1, 0.05
    p.model.sigma2, p.model.lc = 0.2, 0.03
    n = 200000
    parts = kpt.empty_particles(n)
    parts["p"][:] = 1.0
    parts["mu"][:] = np.random.default_rng(1).uniform(-1, 1, n)
    parts["id"][:] = np.arange(n)
    parts["p"][0] = 0.0
    with kpt.Context(n, seed=5) as c:
        c.field_upload(kpt.SLOT_LOWER, 16, xe, ye, planes)
        c.focus_build(p)
        c.upload(parts)
        t, fails = 0.0, 0
        for step in range(5):
            dt, nf = c.focus_step(p, step)
            t += dt
            fails += nf
        got = c.download()
        kperp = c.focus_coefficients(0, nx, ny)[1, 1, 20] * np.sqrt(2.0) ** (1 / 3)
    assert fails == 1 and got["status"][0] == 2 and np.all(got["status"][1:] == 0)
    dy = (got["y"][1:] + 0.5) % 1.0 - 0.5
    assert np.var(dy) == pytest.approx(2 * kperp * t, rel=0.02)
    assert abs(np.mean(dy)) < 4 * np.sqrt(2 * kperp * t / n)
    assert np.all(np.abs(got["mu"]) <= 1.0) and np.all(got["p"][1:] == 1.0)


def test_focus_build_rejects_reference_aborts():
    nx, ny = 16, 8
    xe, ye = np.linspace(-1, 1, nx + 1), np.linspace(-0.5, 0.5, ny + 1)
    one, zero = np.ones((ny, nx)), np.zeros((ny, nx))
    bx = one.copy()
    bx[2, 3] = 0.0
    p = kpt.generic_focus_default_params()
    with kpt.Context(16) as c:
        c.field_upload(kpt.SLOT_LOWER, 16, xe, ye, [bx, zero, zero, 0.1 * one, zero, zero])
        with pytest.raises(kpt.KptError):
            c.focus_build(p)
        with pytest.raises(kpt.KptError):
            c.focus_time_step(p)  # no basis was built
        c.field_upload(kpt.SLOT_LOWER, 16, xe, ye, [one, zero, zero, 0.1 * one, zero, zero])
        bad = p.copy()
        bad.maximum_pitch_angle_cosine = 1.5
        with pytest.raises(kpt.KptError):
            c.focus_build(bad)
        c.focus_build(p)
