"""Two-sample statistics of pushed particle sets (numpy and scipy only).

An FP32 push is compared with the FP64 oracle statistically: the two runs start
from one particle set with independent noise, and their final distributions and
per-particle displacements over the interval are compared with two-sample
chi-square and Kolmogorov-Smirnov tests. Both runs share the start set, so their
final distributions are correlated through it and those tests are lenient; the
displacements, given the start, are independent draws of the dynamics and carry
the sharper test.
"""
from __future__ import annotations

import numpy as np
from scipy import stats


def chi2_two_sample(h1, h2, min_count=5):
    """Two-sample chi-square on histograms of unit-weight counts over the bins
    where either count is >= min_count. Returns (statistic, dof, p-value)."""
    h1 = np.asarray(h1, dtype=np.float64).ravel()
    h2 = np.asarray(h2, dtype=np.float64).ravel()
    n1, n2 = h1.sum(), h2.sum()
    mask = (h1 + h2) >= min_count
    a, b = h1[mask], h2[mask]
    k1 = np.sqrt(n2 / n1)
    k2 = np.sqrt(n1 / n2)
    stat = np.sum((k1 * a - k2 * b) ** 2 / (a + b))
    dof = int(mask.sum()) - 1
    return stat, dof, float(stats.chi2.sf(stat, max(dof, 1)))


def displacements(q, start, x_width, y_width):
    """Per-particle displacements over the interval (minimum image on the periodic
    axes); q and start are aligned particle dicts."""
    def image(d, w):
        return (d + 0.5 * w) % w - 0.5 * w

    out = {"dx": image(q["x"] - start["x"], x_width), "dy": image(q["y"] - start["y"], y_width),
           "dlnp": np.log(q["p"] / start["p"])}
    if "mu" in q:
        out["dmu"] = q["mu"] - start["mu"]
    return out


def displacement_pvalues(want, got, start, x_width, y_width, keys=("dx", "dy", "dlnp", "dmu"), bins=32):
    """KS and chi-square p-values of each displacement; the chi-square histogram
    spans the reference run's 0.5-99.5% range."""
    dw = displacements(want, start, x_width, y_width)
    dg = displacements(got, start, x_width, y_width)
    out = {}
    for k in keys:
        if k not in dw:
            continue
        lo, hi = np.quantile(dw[k], [0.005, 0.995])
        if not hi > lo:
            continue
        out["ks_" + k] = float(stats.ks_2samp(dw[k], dg[k]).pvalue)
        out["chi2_" + k] = chi2_two_sample(np.histogram(dw[k], bins=bins, range=(lo, hi))[0],
                                           np.histogram(dg[k], bins=bins, range=(lo, hi))[0])[2]
    return out


def focused_pvalues(want, got, start, s):
    """Final pitch-angle (32 bins on [-1, 1]), log-p (32 bins on [ln 0.9, ln 1.1e3])
    and spatial (32 x 16) histograms, KS on mu, p and y, and the displacement
    tests of a focused-transport interval."""
    def hists(q):
        hm = np.histogram(q["mu"], bins=32, range=(-1, 1))[0]
        hp = np.histogram(np.log(q["p"]), bins=32, range=(np.log(0.9), np.log(1.1e3)))[0]
        hxy = np.histogram2d(q["y"], q["x"], bins=(16, 32),
                             range=((s.y_lower, s.y_upper), (s.x_lower, s.x_upper)))[0]
        return hm, hp, hxy

    out = {}
    for name, a, b in zip(("chi2_mu", "chi2_logp", "chi2_xy"), hists(want), hists(got)):
        out[name] = chi2_two_sample(a, b)[2]
    for k in ("mu", "p", "y"):
        out["ks_" + k] = float(stats.ks_2samp(want[k], got[k]).pvalue)
    out.update(displacement_pvalues(want, got, start, s.x_width, s.y_width))
    return out


def pathwise_medians(a, b, x_width, y_width, keys=("x", "y", "p", "mu")):
    """Median per-particle differences of two aligned runs: periodic distance in
    x and y, |ln(p_a / p_b)|, |mu_a - mu_b|."""
    out = {}
    for k in keys:
        if k not in a:
            continue
        if k in ("x", "y"):
            w = x_width if k == "x" else y_width
            d = np.abs(a[k] - b[k])
            d = np.minimum(d, w - d)
        elif k == "p":
            d = np.abs(np.log(a[k] / b[k]))
        else:
            d = np.abs(a[k] - b[k])
        out[k] = float(np.median(d))
    return out


def sort_by_id(parts):
    order = np.argsort(parts["id"], kind="stable")
    return {k: v[order] for k, v in parts.items()}
