"""GPU parity of the generic Parker solver (include/ParkerSolver.hpp) against the
oracle: the packed 18-component coefficient basis at every stored node, the
global time step, and ten injected-noise steps to 1e-12, on a uniform grid and
on a grid stretched along both axes (cell centres midway between stored faces,
the cell width of the particle's cell in the time-step bound); with Philox, the
diffusive spread against 2 kappa t."""
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
    p = kpt.generic_default_params()
    p.perpendicular_model = model
    p.perpendicular_parallel_ratio = 0.05 if model == 1 else 0.0
    p.minimum_solar_wind_speed = 1e-3
    p.sigma2, p.lslab, p.l2d = 0.3, 0.05, 0.02
    rng = np.random.default_rng(seed)
    parts = kpt.empty_particles(n)
    parts["x"][:] = rng.uniform(xe[0], xe[-1], n)
    parts["y"][:] = rng.uniform(ye[0], ye[-1], n)
    parts["p"][:] = np.exp(rng.uniform(np.log(0.1), np.log(10.0), n))
    parts["id"][:] = np.arange(n)
    parts["status"][::17] = 2
    return xe, ye, planes, p, parts


@pytest.mark.parametrize("model,stretched", [(0, False), (1, False), (0, True), (1, True)])
def test_generic_basis_and_steps_match_oracle(model, stretched):
    xe, ye, planes, p, parts = _case(model, stretched=stretched)
    n = len(parts["x"])
    of = O.GenericField(xe, ye, planes, p)
    want_c = of.coefficients()
    noise = np.random.default_rng(7).standard_normal((n, 10, 4))
    with kpt.Context(n, seed=11) as c:
        c.field_upload(kpt.SLOT_LOWER, 16, xe, ye, planes)
        c.generic_build(p, kpt.SLOT_LOWER)
        got_c = c.generic_coefficients(kpt.SLOT_LOWER, 64, 32)
        scale = np.maximum(np.abs(want_c).max(axis=(0, 1)), 1e-300)
        err = np.abs(got_c - want_c) / scale
        assert err.max() < 1e-12, (err.max(), np.unravel_index(err.argmax(), err.shape))
        c.upload(parts)
        assert c.generic_time_step(p) == pytest.approx(of.time_step(parts), rel=1e-12)
        c.set_noise(noise)
        want = {k: v.copy() for k, v in parts.items()}
        for step in range(10):
            dt, nf = c.generic_step(p, step)
            dto, nfo = of.step(want, step, noise=noise)
            assert nf == nfo == 0
            assert dt == pytest.approx(dto, rel=1e-12)
        got = c.download()
    for k, w in (("x", xe[-1] - xe[0]), ("y", ye[-1] - ye[0])):
        d = np.abs(got[k] - want[k])
        d = np.minimum(d, w - d) / w
        assert d.max() < 1e-12, k
    np.testing.assert_allclose(got["p"], want["p"], rtol=1e-12)
    np.testing.assert_array_equal(got["status"], want["status"])


def test_generic_philox_spread_matches_diffusion():
    """Uniform B along x, uniform V: after k steps the field-aligned spread of x
    (drift removed) has variance 2 kappa_par t and y has 2 kappa_perp t."""
    nx, ny = 32, 16
    xe, ye = np.linspace(-1, 1, nx + 1), np.linspace(-0.5, 0.5, ny + 1)
    one, zero = np.ones((ny, nx)), np.zeros((ny, nx))
    planes = [one, zero, zero, 0.3 * one, zero, zero]
    p = kpt.generic_default_params()
    p.perpendicular_model, p.perpendicular_parallel_ratio = 1, 0.25
    n = 200000
    rng = np.random.default_rng(2)
    parts = kpt.empty_particles(n)
    parts["x"][:] = 0.0
    parts["y"][:] = 0.0
    parts["p"][:] = 1.0
    parts["id"][:] = np.arange(n)
    with kpt.Context(n, seed=5) as c:
        c.field_upload(kpt.SLOT_LOWER, 16, xe, ye, planes)
        c.generic_build(p)
        c.upload(parts)
        t = 0.0
        for step in range(5):
            dt, nf = c.generic_step(p, step)
            assert nf == 0
            t += dt
        got = c.download()
        kpar1 = c.generic_coefficients(0, nx, ny)[1, 1, 15]
    kpar = kpar1 * np.sqrt(2.0) ** (1 / 3)
    dx = got["x"] - 0.3 * t
    dx = (dx + 1.0) % 2.0 - 1.0
    dy = (got["y"] + 0.5) % 1.0 - 0.5
    assert np.var(dx) == pytest.approx(2 * kpar * t, rel=0.02)
    assert np.var(dy) == pytest.approx(2 * 0.25 * kpar * t, rel=0.02)
    assert abs(np.mean(dy)) < 4 * np.sqrt(2 * 0.25 * kpar * t / n)
