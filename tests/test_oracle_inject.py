"""Pins the oracle's generic particle generator (include/ParticleGenerator.hpp)
without a GPU: the field-threshold candidate set equals a brute-force
enumeration of the reference predicate over the cell centres (every comparison,
ties included, inclusive regions) with the running sum of cell areas; the grid
set's total is the domain area and its uniform-grid weights are k dx dy; a draw
picks the first candidate whose cumulative weight exceeds total * u (Python's
bisect_right on an independent Philox restatement); volume weighting with
uniform in-cell draws gives a uniform density on a stretched grid; rejection
sampling fails with probability (1 - a)^m; recycling touches only the selected
statuses and keeps ids."""
import bisect
import math

import numpy as np
import pytest
from scipy import stats

import kpt
import kpt_oracle as O
import workloads as W

TAG_INJECT = 0x494E0000


def _field(nx=12, ny=10, stretched=False, seed=3):
    xe = np.linspace(-1.0, 1.0, nx + 1)
    ye = W.stretched_edges(ny, -0.5, 0.5, 6.0) if stretched else np.linspace(-0.5, 0.5, ny + 1)
    rng = np.random.default_rng(seed)
    planes = [rng.normal(0.0, 1.0, (ny, nx)) for _ in range(6)]
    planes[2] = np.round(planes[2], 1)  # ties with thresholds on a 0.1 grid
    return xe, ye, planes


def _centres(xe, ye):
    nx, ny = len(xe) - 1, len(ye) - 1
    uniform = [np.allclose(np.diff(e), np.diff(e)[0], rtol=1e-13, atol=0) for e in (xe, ye)]
    cx = [xe[0] + (i + 0.5) * (xe[-1] - xe[0]) / nx if uniform[0] else 0.5 * (xe[i] + xe[i + 1]) for i in range(nx)]
    cy = [ye[0] + (j + 0.5) * (ye[-1] - ye[0]) / ny if uniform[1] else 0.5 * (ye[j] + ye[j + 1]) for j in range(ny)]
    return cx, cy


def _passes(v, thr, cmp):
    return [v > thr, v >= thr, v < thr, v <= thr, abs(v) > thr, abs(v) < thr][cmp]


@pytest.mark.parametrize("stretched", [False, True])
@pytest.mark.parametrize("cmp", range(6))
def test_field_candidates_equal_brute_force(oracle_mod, stretched, cmp):
    xe, ye, planes = _field(stretched=stretched)
    nx, ny = len(xe) - 1, len(ye) - 1
    thr = 0.3 if cmp < 4 else 0.5
    region = (-0.55, 0.77, -0.33, 0.47)  # bounds off the cell centres (a bound on a centre is a rounding tie)
    cells, cum, total = O.candidates_field(planes, xe, ye, (2, cmp, thr), region)
    cx, cy = _centres(xe, ye)
    want, areas = [], []
    for j in range(ny):
        for i in range(nx):
            if not (region[0] <= cx[i] <= region[1] and region[2] <= cy[j] <= region[3]):
                continue
            if _passes(float(planes[2][j, i]), thr, cmp):
                want.append(i + nx * j)
                areas.append((xe[i + 1] - xe[i]) * (ye[j + 1] - ye[j]))
    assert cells.tolist() == want
    np.testing.assert_allclose(cum, np.cumsum(areas), rtol=1e-12)
    assert total == pytest.approx(sum(areas), rel=1e-12)


def test_ties_are_split_by_the_comparison(oracle_mod):
    xe, ye, planes = _field()
    n_eq = int(np.sum(planes[2] == 0.3))
    assert n_eq > 0
    counts = [len(O.candidates_field(planes, xe, ye, (2, c, 0.3))[0]) for c in range(4)]
    n = planes[2].size
    assert counts[1] - counts[0] == n_eq            # >= minus >
    assert counts[3] - counts[2] == n_eq            # <= minus <
    assert counts[0] + counts[3] == n and counts[1] + counts[2] == n


@pytest.mark.parametrize("stretched", [False, True])
def test_grid_candidates_total_is_the_domain_area(oracle_mod, stretched):
    xe, ye, _ = _field(16, 9, stretched)
    cum, total = O.candidates_grid(xe, ye)
    assert len(cum) == 16 * 9 and np.all(np.diff(cum) > 0)
    assert total == pytest.approx((xe[-1] - xe[0]) * (ye[-1] - ye[0]), rel=1e-13)
    if not stretched:
        np.testing.assert_allclose(cum, np.arange(1, 16 * 9 + 1) * (2.0 / 16) * (1.0 / 9), rtol=1e-13)


def _kin(**kw):
    return kpt.injection_kinematics(**{"momentum": 2.0, "mu": 0.25, "weight": 0.5, "status": 0, "time": 0.3,
                                       "step_dt": 1e-3, **kw})


def test_candidate_draw_is_the_first_cumulative_weight_above_the_target(oracle_mod):
    xe, ye, planes = _field(20, 14, True)
    cells, cum, total = O.candidates_field(planes, xe, ye, (0, kpt.CMP_GREATER, 0.0))
    n, seed, call, root0 = 3000, 77, 5, 1000
    parts = kpt.empty_particles(n)
    O.generator_inject_candidates(0, cells, cum, total, xe, ye, n, _kin(), 1.0, seed, call, root0, parts, 0)
    cx, cy = _centres(xe, ye)
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for m in range(0, n, 7):
        root = root0 + m
        r = O.philox4x32_10((root & 0xFFFFFFFF, TAG_INJECT, call, root >> 32), key)
        u = (r[0] + 0.5) / 4294967296.0
        idx = bisect.bisect_right(cum.tolist(), total * u)
        c = int(cells[idx])
        assert parts["x"][m] == pytest.approx(cx[c % 20], abs=1e-15)
        assert parts["y"][m] == pytest.approx(cy[c // 20], abs=1e-15)
        assert parts["id"][m] == root
    assert np.all(parts["p"] == 2.0) and np.all(parts["mu"] == 0.25) and np.all(parts["weight"] == 0.5)
    assert np.all(parts["t"] == 0.3) and np.all(parts["step_dt"] == 1e-3) and np.all(parts["status"] == 0)
    # counts per candidate follow the areas
    ci = np.searchsorted(xe, parts["x"]) - 1  # centres lie strictly inside their cells
    cj = np.searchsorted(ye, parts["y"]) - 1
    hit = np.bincount(ci + 20 * cj, minlength=20 * 14)[cells]
    assert hit.sum() == n
    w = np.diff(np.concatenate([[0.0], cum])) / total
    assert stats.chisquare(hit, w * n).pvalue > 0.01


def test_grid_draws_are_uniform_in_space_on_a_stretched_grid(oracle_mod):
    xe, ye, _ = _field(8, 32, True)
    cum, total = O.candidates_grid(xe, ye)
    n = 40000
    parts = kpt.empty_particles(n)
    O.generator_inject_candidates(1, None, cum, total, xe, ye, n, _kin(), 1.0, 9, 0, 0, parts, 0)
    assert np.all((parts["x"] >= -1.0) & (parts["x"] <= 1.0) & (parts["y"] >= -0.5) & (parts["y"] <= 0.5))
    assert stats.kstest(parts["y"], stats.uniform(-0.5, 1.0).cdf).pvalue > 0.01
    assert stats.kstest(parts["x"], stats.uniform(-1.0, 2.0).cdf).pvalue > 0.01


@pytest.mark.parametrize("attempts", [1, 3])
def test_rejection_sampling_fails_with_probability_one_minus_a_to_the_m(oracle_mod, attempts):
    xe, ye = np.linspace(0.0, 1.0, 11), np.linspace(0.0, 1.0, 11)
    one = np.ones((10, 10))
    bz = np.where(np.arange(10)[None, :] < 3, 1.0, -1.0) * one  # Bz > 0 on x < 0.3
    planes = [one, one, bz, one, one, one]
    n = 20000
    parts = kpt.empty_particles(n)
    rej = O.generator_inject_uniform(planes, xe, ye, None, (2, kpt.CMP_GREATER, 0.0), attempts, n, _kin(), 1.0, 4, 0,
                                     0, parts, 0)
    ok = parts["status"] == 0
    assert rej == int(np.sum(~ok)) and np.all(parts["status"][~ok] == 2)
    assert np.all(parts["x"][ok] < 0.3)
    p_fail = 0.7 ** attempts
    assert abs(rej / n - p_fail) < 5 * math.sqrt(p_fail * (1 - p_fail) / n)
    # without a predicate every draw is accepted, uniform in the region
    parts2 = kpt.empty_particles(n)
    assert O.generator_inject_uniform(planes, xe, ye, (0.2, 0.4, 0.1, 0.9), None, 1, n, _kin(), 1.0, 4, 0, 0,
                                      parts2, 0) == 0
    assert np.all((parts2["x"] >= 0.2) & (parts2["x"] < 0.4) & (parts2["y"] >= 0.1) & (parts2["y"] < 0.9))


@pytest.mark.parametrize("selection,hit", [(kpt.RECYCLE_ESCAPED, (1,)), (kpt.RECYCLE_INACTIVE, (2,)),
                                           (kpt.RECYCLE_ANY, (1, 2))])
def test_recycling_touches_only_the_selected_statuses(oracle_mod, selection, hit):
    xe, ye, planes = _field()
    n = 3000
    rng = np.random.default_rng(2)
    parts = kpt.empty_particles(n)
    parts["status"][:] = rng.integers(0, 3, n)
    parts["id"][:] = rng.permutation(n).astype(np.uint64) + 10
    parts["x"][:] = 5.0
    before = {k: v.copy() for k, v in parts.items()}
    cum, total = O.candidates_grid(xe, ye)
    rec, rej = O.generator_recycle(0, _kin(status=0), 1.0, 3, 2, parts, n, selection, xe, ye, kind=1, cum=cum,
                                   total=total)
    sel = np.isin(before["status"], hit)
    assert rec == int(sel.sum()) and rej == 0
    np.testing.assert_array_equal(parts["id"], before["id"])
    assert np.all(parts["status"][sel] == 0) and np.all(np.abs(parts["x"][sel]) <= 1.0)
    for k in ("x", "y", "status", "p"):
        np.testing.assert_array_equal(parts[k][~sel], before[k][~sel])
