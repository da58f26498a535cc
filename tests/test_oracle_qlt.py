"""Pins the oracle's slab-QLT pitch-angle closure (include/LegencyModel.hpp:539-689),
the reflecting mu boundary (include/FocusSolver.hpp:454-470) and the Picard mu
update (:700-719) with closed forms, symmetries, finite differences and an
independent Picard loop. No GPU is needed."""
import numpy as np
import pytest

import kpt
import kpt_oracle as O
from helpers import const_field, copy_parts, random_particles, uniform_edges


def test_qlt_gamma2_closed_form(oracle_mod):
    """gamma_k = 2: C_2 = 2 sin(pi/2)/pi = 2/pi, so
    D = (1/2) Omega sigma2 (1 - mu^2) xi mu_h / (mu_h^2 + xi^2)."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        mu = rng.uniform(-0.999, 0.999)
        om, v, s2, lc = rng.uniform(0.1, 5), rng.uniform(0.2, 1), rng.uniform(0.01, 1), rng.uniform(0.005, 0.5)
        D, _ = O.qlt_scattering(mu, om, v, s2, lc, 2.0, 1e-6)
        xi = om * lc / v
        mh = abs(mu) + 1e-6
        assert D == pytest.approx(0.5 * om * s2 * (1 - mu * mu) * xi * mh / (mh * mh + xi * xi), rel=1e-13)


@pytest.mark.parametrize("gk", [1.5, 5.0 / 3.0, 2.0])
def test_qlt_derivative_matches_finite_difference(oracle_mod, gk):
    for mu in (-0.93, -0.4, -0.05, 0.02, 0.3, 0.77):
        args = (2.3, 0.8, 0.1, 0.02, gk, 1e-6)
        D, dD = O.qlt_scattering(mu, *args)
        eps = 1e-6
        fd = (O.qlt_scattering(mu + eps, *args)[0] - O.qlt_scattering(mu - eps, *args)[0]) / (2 * eps)
        assert dD == pytest.approx(fd, rel=1e-6)


def test_qlt_symmetries_and_limits(oracle_mod):
    args = (1.7, 0.9, 0.2, 0.03, 5.0 / 3.0, 1e-6)
    for mu in (0.1, 0.5, 0.9):
        Dp, dp = O.qlt_scattering(mu, *args)
        Dm, dm = O.qlt_scattering(-mu, *args)
        assert Dp == Dm and dp == -dm
    assert O.qlt_scattering(1.0, *args) == (0.0, 0.0)
    assert O.qlt_scattering(-1.0, *args) == (0.0, 0.0)
    assert O.qlt_scattering(0.0, *args)[1] == 0.0  # sgn(0) = 0
    D1 = O.qlt_scattering(0.4, 1.7, 0.9, 0.2, 0.03, 5 / 3, 1e-6)[0]
    D2 = O.qlt_scattering(0.4, 1.7, 0.9, 0.6, 0.03, 5 / 3, 1e-6)[0]
    assert D2 == pytest.approx(3 * D1, rel=1e-14)  # linear in sigma2
    # Omega and v scaled together keep xi: D scales with Omega
    D3 = O.qlt_scattering(0.4, 3.4, 1.8, 0.2, 0.03, 5 / 3, 1e-6)[0]
    assert D3 == pytest.approx(2 * D1, rel=1e-14)


def test_reflection_identities(oracle_mod):
    assert O.reflect_mu(1.0 + 0.25, 1.0) == pytest.approx(0.75)
    assert O.reflect_mu(-1.0 - 0.25, 1.0) == pytest.approx(-0.75)
    assert O.reflect_mu(3.5, 1.0) == pytest.approx(-0.5)  # 3.5 -> -1.5 -> -0.5
    assert O.reflect_mu(0.3, 0.9) == 0.3
    assert O.reflect_mu(0.95, 0.9) == pytest.approx(0.85)
    for mu in np.linspace(-50, 50, 1001):  # period 4 upper, range [-upper, upper]
        r = O.reflect_mu(mu, 1.0)
        assert -1.0 <= r <= 1.0
        assert r == pytest.approx(O.reflect_mu(mu + 4.0, 1.0), abs=1e-9)


def _qlt_settings(xe, ye, **kw):
    h = kpt.default_host_settings()
    h.focused_transport = 1
    h.magnetic_dependency = 0
    h.momentum_dependency = 0
    h.p0 = 1.0
    h.kpara0 = 1e-6
    h.kperp_over_kpara = 0.0
    h.mhd_dt = 1e-3
    h.dt_min_rel = 1.0
    h.dt_max_rel = 1.0
    h.mu_max = 1.0
    s = kpt.make_device_settings(h, xe, ye)
    s.closure = kpt.CLOSURE_QLT
    s.qlt_sigma2 = 0.1
    s.qlt_correlation_length = 0.02
    s.max_particle_steps_per_interval = 1
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_focused_qlt_step_matches_independent_picard(oracle_mod):
    """Uniform B = (b0, 0, 0), V = 0: one step of the QLT closure moves x by
    v mu dt with v = p / sqrt(1 + p^2), keeps p, and updates mu by three Picard
    passes mu_{k+1} = R(mu_0 + dt (dD/dmu)(mu_k) + sqrt(2 D(mu_k) dt) r), R the
    reflection, the focusing terms being zero for a uniform field."""
    xe, ye = uniform_edges(16, -1, 1), uniform_edges(16, -1, 1)
    b0 = 1.5
    lo = O.Field(xe, ye, const_field(16, 16, bx=b0), 21)
    s = _qlt_settings(xe, ye)
    mu0 = np.array([0.6, -0.2, 0.97, -0.995, 0.01])
    n = len(mu0)
    parts = random_particles(n, s, np.random.default_rng(0))
    parts["x"][:] = 0.0
    parts["y"][:] = 0.0
    parts["p"][:] = [0.5, 2.0, 10.0, 1.0, 3.0]
    parts["mu"][:] = mu0
    noise = np.zeros((n, 1, 4))
    noise[:, 0, 3] = [0.7, -1.2, 1.5, -1.6, 0.3]
    start = copy_parts(parts)
    O.push_interval(lo, lo, s, 0.0, 0, 1, parts, noise=noise)
    dt = s.dt_min
    p = start["p"]
    gamma = np.sqrt(1 + p * p)
    v = p / gamma
    np.testing.assert_allclose(parts["x"], v * mu0 * dt, rtol=1e-12)
    np.testing.assert_array_equal(parts["p"], p)
    omega = np.abs(s.charge) * b0 / gamma
    for i in range(n):
        trial = mu0[i]
        for _ in range(3):
            D, dD = O.qlt_scattering(O.reflect_mu(trial), omega[i], v[i], s.qlt_sigma2,
                                     s.qlt_correlation_length, s.qlt_spectral_index, s.qlt_h0)
            trial = O.reflect_mu(mu0[i] + dt * dD + np.sqrt(2 * D * dt) * noise[i, 0, 3])
        assert parts["mu"][i] == pytest.approx(trial, abs=1e-14), i


def test_focused_qlt_picard_stationary_density(oracle_mod):
    """Pure pitch-angle scattering (uniform B, V = 0). The Picard update evaluates
    D at the end point of the step with the same noise, so it integrates
    d mu = D' dt + sqrt(2 D) dW in the backward (Haenggi-Klimontovich) sense:
    the Ito drift becomes 2 D', and the stationary solution of
    d/dmu (-D' f + D f') = 0 is f ~ D(mu), not the uniform density. A population
    drawn from f ~ D stays so (KS p > 0.01); a uniform one does not."""
    from scipy import stats

    xe, ye = uniform_edges(16, -1, 1), uniform_edges(16, -1, 1)
    lo = O.Field(xe, ye, const_field(16, 16, bx=1.0), 21)
    s = _qlt_settings(xe, ye, qlt_sigma2=0.5, qlt_correlation_length=0.5)
    s.max_particle_steps_per_interval = 200000
    s.mhd_dt = 0.4
    s.dt_min = s.dt_max = 4e-4
    gamma_p = np.sqrt(2.0)
    omega, v = 1.0 / gamma_p, 1.0 / gamma_p
    grid = np.linspace(-1, 1, 20001)
    dens = np.array([O.qlt_scattering(m, omega, v, 0.5, 0.5)[0] for m in grid])
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (dens[1:] + dens[:-1]) * np.diff(grid))])
    cdf /= cdf[-1]
    d_cdf = lambda m: np.interp(m, grid, cdf)
    n = 20000
    rng = np.random.default_rng(3)
    parts = random_particles(n, s, rng)
    parts["x"][:] = 0.0
    parts["y"][:] = 0.0
    parts["p"][:] = 1.0
    parts["mu"][:] = np.interp(rng.uniform(size=n), cdf, grid)
    O.push_interval(lo, lo, s, 0.0, 0, 5, parts)
    assert np.all(np.abs(parts["mu"]) <= 1.0)
    assert stats.kstest(parts["mu"], d_cdf).pvalue > 0.01
    parts["mu"][:] = rng.uniform(-1, 1, n)
    parts["t"][:] = 0.0
    O.push_interval(lo, lo, s, 0.0, 0, 6, parts)
    assert stats.kstest(parts["mu"], stats.uniform(loc=-1, scale=2).cdf).pvalue < 1e-3
