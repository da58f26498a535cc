"""Pins the CPU oracle of the learned pitch-angle closure (oracle/kpt_mlp_oracle.py):
the reparametrised targets against the slab-QLT coefficient of the oracle
(include/LegencyModel.hpp:539-689), the analytic training gradient against
central finite differences (both activations), the hard-swish activation
against its closed form, and the BF16 restatement against plain float64.
No GPU is needed."""
import numpy as np
import pytest

import kpt
import kpt_mlp_oracle as M
import kpt_oracle as O

HULL = kpt.mlp_default_train_opts().hull.as_dict()


def test_targets_reproduce_slab_qlt_coefficient(oracle_mod):
    """D = pref (1 - mu^2) exp(ln R) and the closure's dD/dmu from y1 equal the
    oracle's qlt_scattering for random (mu, Omega, v, sigma2, Lc, gamma_k)."""
    rng = np.random.default_rng(1)
    hull = dict(HULL, out_mean=0.0, out_scale=1.0)
    for _ in range(300):
        mu = rng.uniform(0.02, 0.999) * rng.choice([-1, 1])
        om, v, s2, lc = rng.uniform(0.05, 3), rng.uniform(0.3, 1), rng.uniform(0.01, 1), rng.uniform(1e-4, 0.1)
        gk = rng.uniform(1.5, 2.0)
        D, dD = O.qlt_scattering(mu, om, v, s2, lc, gk, 1e-6)
        xi = om * lc / v
        ln_r, y1 = M.exact_targets(abs(mu), np.log(xi), gk)
        pref = 0.25 * np.pi * om * s2 * gk * np.sin(np.pi / gk) / np.pi
        Dm, dDm = M.closure(ln_r, y1, mu, pref, hull)
        assert Dm == pytest.approx(D, rel=1e-11)
        assert dDm == pytest.approx(dD, rel=1e-9, abs=1e-12 * abs(D))


def test_features_cover_the_hull():
    f = M.features(np.array([HULL["mu_lo"], 1.0]), np.array([HULL["log_xi_lo"], HULL["log_xi_hi"]]),
                   np.array([HULL["gamma_lo"], HULL["gamma_hi"]]), HULL)
    np.testing.assert_allclose(f, [[-1, -1, -1], [1, 1, 1]], atol=1e-12)


def _random_net(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM)
    p[M.OFF_W1:M.OFF_B1] = rng.standard_normal(M.OFF_B1 - M.OFF_W1) * np.sqrt(1 / 3) * scale
    p[M.OFF_B1:M.OFF_W2] = rng.standard_normal(64) * 0.1
    for a, b in ((M.OFF_W2, M.OFF_B2), (M.OFF_W3, M.OFF_B3), (M.OFF_W4, M.OFF_B4)):
        p[a:b] = rng.standard_normal(b - a) * np.sqrt(1 / 64) * scale
    for a, b in ((M.OFF_B2, M.OFF_W3), (M.OFF_B3, M.OFF_W4), (M.OFF_B4, M.NPARAM)):
        p[a:b] = rng.standard_normal(b - a) * 0.1
    return p


@pytest.mark.parametrize("act", [M.ACT_SILU, M.ACT_HARD_SWISH])
def test_gradient_matches_finite_differences(act):
    rng = np.random.default_rng(2)
    p = _random_net()
    x = rng.uniform(-1, 1, (37, 3))
    y = rng.standard_normal((37, 2))
    _, g = M.loss_and_grad(p, x, y, act)
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in
                          ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2),
                           (M.OFF_B2, M.OFF_W3), (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.OFF_W4),
                           (M.OFF_W4, M.OFF_B4))] + [np.arange(M.OFF_B4, M.NPARAM)])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_and_grad(pp, x, y, act)[0] - M.loss_and_grad(pm, x, y, act)[0]) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i  # central-difference floor ~1e-10


def test_hard_swish_closed_form():
    """z max(0, min(1, (z + 3) / 6)): zero below -3, identity above 3, the
    parabola z (z + 3) / 6 between; continuous with a derivative that jumps at
    the ends of the ramp (0 -> -1/2 at -3, 3/2 -> 1 at 3)."""
    z = np.array([-7.0, -3.0, -1.5, 0.0, 1.0, 3.0, 4.5])
    np.testing.assert_allclose(M.hard_swish(z), [0.0, 0.0, -0.375, 0.0, 2.0 / 3.0, 3.0, 4.5], atol=1e-15)
    zz = np.concatenate([np.linspace(-6, -3.01, 50), np.linspace(-2.99, 2.99, 200), np.linspace(3.01, 6, 50)])
    fd = (M.hard_swish(zz + 1e-7) - M.hard_swish(zz - 1e-7)) / 2e-7
    np.testing.assert_allclose(M.dhard_swish(zz), fd, atol=1e-7)
    assert M.dhard_swish(np.array([-3.0 + 1e-9]))[0] == pytest.approx(-0.5)
    assert M.dhard_swish(np.array([3.0 - 1e-9]))[0] == pytest.approx(1.5)
    # the float32 form of the kernels rounds, but agrees to float32 precision
    zf = np.random.default_rng(0).uniform(-8, 8, 4096).astype(np.float32)
    np.testing.assert_allclose(M._hard_swish_f32(zf), M.hard_swish(zf.astype(np.float64)), rtol=3e-7, atol=3e-7)


@pytest.mark.parametrize("act", [M.ACT_SILU, M.ACT_HARD_SWISH])
def test_bf16_restatement_is_close_to_float64(act):
    p = _random_net(3)
    x = np.random.default_rng(4).uniform(-1, 1, (256, 3))
    _, y = M.forward(p, x, act)
    yb = M.forward_bf16(p, x, act)
    assert np.max(np.abs(yb - y)) < 0.03 * max(1.0, np.max(np.abs(y)))
    assert np.max(np.abs(yb - y)) > 0  # BF16 rounding is visible
    # bf16() rounds to nearest even on the 16-bit boundary
    assert M.bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert M.bf16(np.float32(1.0 + 3 * 2 ** -9)) == 1.0 + 2 ** -7
