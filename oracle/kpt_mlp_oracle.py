"""CPU oracle of the learned pitch-angle closure (numpy, float64).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the cpu_baseline leg of
bench.py as the checker; the product path never imports it.

The closure stands in for the resonance part of the slab-QLT coefficient of
LegencyModel::evaluate_pitch_angle_scattering (include/LegencyModel.hpp:539-689):
    D = (pi/4) Omega sigma2 C_gamma (1 - mu^2) R(|mu|, xi, gamma_k),
    R = xi mu_h^(gamma_k - 1) / (mu_h^gamma_k + xi^gamma_k), mu_h = |mu| + h0.
The network maps scaled (ln|mu|, ln xi, gamma_k) to
y0 = (ln R - out_mean) / out_scale and y1 = |mu| d ln R / d|mu| through
3 -> 64 -> 64 -> 64 -> 2 with SiLU or hard-swish. This module restates the targets, the
features, the forward pass, the analytic gradient of the training loss
L = 1/(2n) sum (yhat - y)^2 and the mapping from outputs back to (D, dD/dmu).
The gradient is pinned by central finite differences and the targets by the
slab-QLT coefficient of the oracle (tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np

IN, H, OUT = 3, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + OUT * H
NPARAM = OFF_B4 + OUT


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    assert p.size == NPARAM
    return (p[OFF_W1:OFF_B1].reshape(H, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(H, H), p[OFF_B3:OFF_W4],
            p[OFF_W4:OFF_B4].reshape(OUT, H), p[OFF_B4:])


ACT_SILU, ACT_HARD_SWISH = 0, 1   # hidden-layer activations (KPT_MLP_SILU, KPT_MLP_HARD_SWISH)
SIXTH = np.float32(1.0 / 6.0)      # the float constant of the kernels' hard-swish ramp


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def hard_swish(z):
    """z u, u = clamp(z / 6 + 1/2, 0, 1)."""
    return z * np.clip(z / 6.0 + 0.5, 0.0, 1.0)


def dhard_swish(z):
    """u + z / 6 = z / 3 + 1/2 on the ramp |z| < 3, 0 below it, 1 above it."""
    return np.where(z <= -3.0, 0.0, np.where(z >= 3.0, 1.0, z / 3.0 + 0.5))


def activation(act):
    """(f, f') of a hidden-layer activation code."""
    if act == ACT_SILU:
        return silu, dsilu
    if act == ACT_HARD_SWISH:
        return hard_swish, dhard_swish
    raise ValueError(f"unknown activation {act}")


def exact_targets(mu_abs, log_xi, gk, h0=1e-6):
    """ln R and |mu| d ln R / d|mu| (float64)."""
    mh = np.asarray(mu_abs, dtype=np.float64) + h0
    xi = np.exp(log_xi)
    mg, xg = mh ** gk, xi ** gk
    ln_r = log_xi + (gk - 1.0) * np.log(mh) - np.log(mg + xg)
    y1 = mu_abs * ((gk - 1.0) / mh - gk * mh ** (gk - 1.0) / (mg + xg))
    return ln_r, y1


def features(mu_abs, log_xi, gk, hull: dict):
    """Affine scaling of (ln|mu|, ln xi, gamma_k) onto [-1, 1] over the hull."""
    lmu = np.log(hull["mu_lo"])
    f0 = (np.log(mu_abs) - 0.5 * lmu) * (-2.0 / lmu)
    f1 = (log_xi - 0.5 * (hull["log_xi_lo"] + hull["log_xi_hi"])) * (2.0 / (hull["log_xi_hi"] - hull["log_xi_lo"]))
    f2 = (gk - 0.5 * (hull["gamma_lo"] + hull["gamma_hi"])) * (2.0 / (hull["gamma_hi"] - hull["gamma_lo"]))
    return np.stack(np.broadcast_arrays(f0, f1, f2), axis=-1)


def closure(y0, y1, mu, pref, hull: dict):
    """(D, dD/dmu) from the network outputs: D = P (1 - mu^2) R,
    dD/dmu = P R ((1 - mu^2) sgn(mu) y1 / |mu| - 2 mu)."""
    r = np.exp(hull["out_mean"] + hull["out_scale"] * np.asarray(y0, dtype=np.float64))
    omm = 1.0 - mu * mu
    return pref * omm * r, pref * r * (omm * np.sign(mu) * y1 / np.abs(mu) - 2.0 * mu)


def forward(p, x, act=ACT_SILU):
    """Pre-activations (z1, z2, z3) and outputs for rows of x (n x 3)."""
    f, _ = activation(act)
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ w1.T + b1
    z2 = f(z1) @ w2.T + b2
    z3 = f(z2) @ w3.T + b3
    return (z1, z2, z3), f(z3) @ w4.T + b4


def loss_and_grad(p, x, y, act=ACT_SILU):
    """L = 1/(2n) sum_b sum_o (yhat - y)^2 and dL/dp (float64, analytic)."""
    f, df = activation(act)
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = x.shape[0]
    (z1, z2, z3), yhat = forward(p, x, act)
    e = yhat - y
    loss = 0.5 * np.sum(e * e) / n
    dy = e / n
    h1, h2, h3 = f(z1), f(z2), f(z3)
    g = np.zeros(NPARAM)
    g[OFF_W4:OFF_B4] = (dy.T @ h3).ravel()
    g[OFF_B4:] = dy.sum(0)
    d3 = (dy @ w4) * df(z3)
    g[OFF_W3:OFF_B3] = (d3.T @ h2).ravel()
    g[OFF_B3:OFF_W4] = d3.sum(0)
    d2 = (d3 @ w3) * df(z2)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = (d2 @ w2) * df(z1)
    g[OFF_W1:OFF_B1] = (d1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    return loss, g


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


FOLD_PRE = np.float32(-1.4426950408889634)   # -log2(e)
FOLD_OUT = np.float32(-0.6931471805599453)   # -ln(2)


def _scaled(w, c):
    """FP32 product with the fold constant (as the image builder forms it)."""
    return np.asarray(w, dtype=np.float32) * c


def _hard_swish_f32(z):
    """The kernels' hard-swish: u = fma(z, float(1/6), 1/2) rounded once to
    float32 and clamped to [0, 1], then z u rounded to float32."""
    z = np.asarray(z, dtype=np.float32)
    u = (z.astype(np.float64) * np.float64(SIXTH) + 0.5).astype(np.float32)
    return (z * np.clip(u, np.float32(0.0), np.float32(1.0))).astype(np.float32)


def forward_bf16(p, x, act=ACT_SILU):
    """The wavefront MFMA evaluation restated: BF16 weights, biases of layer 1
    in BF16 (they ride in the K column), BF16 inputs and hidden activations,
    FP32-or-better accumulation, FP32 biases of layers 2-4. For SiLU the image
    folds the exp2 scale into the network: pre-activations are z' = -log2(e) z,
    activations h' = z' / (1 + 2^z') = -log2(e) silu(z), so layer 1 and the
    biases of layers 2 and 3 carry -log2(e), layers 2 and 3 keep their weights
    and layer 4 carries -ln(2); in exact arithmetic the function is unchanged.
    A hard-swish network is stored as trained."""
    w1, b1, w2, b2, w3, b3, w4, b4 = unpack(p)
    if act == ACT_SILU:
        pre, out = FOLD_PRE, FOLD_OUT
        f = lambda zs: bf16(zs / (1.0 + np.exp2(zs)))  # noqa: E731
    elif act == ACT_HARD_SWISH:
        pre, out = np.float32(1.0), np.float32(1.0)
        f = lambda z: bf16(_hard_swish_f32(z))  # noqa: E731
    else:
        raise ValueError(f"unknown activation {act}")
    x = bf16(x)
    z1 = x @ bf16(_scaled(w1, pre)).T + bf16(_scaled(b1, pre))
    h = f(z1)
    z2 = h @ bf16(w2).T + _scaled(b2, pre)
    h = f(z2)
    z3 = h @ bf16(w3).T + _scaled(b3, pre)
    h = f(z3)
    return h @ bf16(_scaled(w4, out)).T + np.float32(b4)
