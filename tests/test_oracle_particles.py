"""Pins the oracle's ParticleSystem lifecycle passes (include/ParticleSystem.hpp:
split_particles_on_energy_growth, reset_split_baselines,
apply_grid_boundary_conditions) with closed forms and invariants: repeated
energy splits of a parent whose energy grew by r^T end in 2^T equal-weight
descendants at level T with the weight sum unchanged; the relativistic kinetic
energy reduces to p^2/2m and pc in its limits; wrapped coordinates land in the
domain and differ by whole periods; escape is decided on the closed extent.
No GPU is needed."""
import numpy as np
import pytest

import kpt
import kpt_oracle as O


def _props(m=1.0, c=1.0):
    p = kpt.default_particle_properties()
    p.rest_mass, p.speed_of_light = m, c
    return p


def _policy(ratio=2.0, wmin=0.0):
    p = kpt.default_energy_split_policy()
    p.energy_split_ratio, p.minimum_child_weight = ratio, wmin
    return p


def _ek(p, m=1.0, c=1.0):
    return (np.sqrt(1.0 + (p / (m * c)) ** 2) - 1.0) * m * c * c


def _set(n, cap, p, e0, weight=1.0):
    parts = kpt.empty_particles(cap)
    parts["status"][:] = 2
    parts["status"][:n] = 0
    parts["p"][:n] = p
    parts["e0"][:n] = e0
    parts["weight"][:n] = weight
    parts["id"][:n] = np.arange(n)
    return parts


def test_defaults_match_reference():
    pr, po = kpt.default_particle_properties(), kpt.default_energy_split_policy()
    assert (pr.rest_mass, pr.charge, pr.speed_of_light, pr.energy_scale_erg) == (1.0, 1.0, 1.0, 1.0)
    assert (po.energy_split_ratio, po.minimum_child_weight) == (2.0, 0.0)


@pytest.mark.parametrize("m,c", [(1.0, 1.0), (2.0, 3.0)])
def test_kinetic_energy_limits_through_reset(m, c):
    p = np.array([1e-4, 1e-3, 1.0, 1e3, 1e6]) * m * c
    parts = _set(len(p), len(p), p, 0.0)
    parts["split_level"][:] = 7
    O.reset_split_baselines(_props(m, c), parts, len(p))
    e = parts["e0"]
    assert np.all(parts["split_level"] == 0)
    np.testing.assert_allclose(e[:2], p[:2] ** 2 / (2 * m), rtol=1e-6)          # non-relativistic
    np.testing.assert_allclose(e[-1], p[-1] * c - m * c * c, rtol=1e-9)          # ultra-relativistic
    np.testing.assert_allclose(e[2], (np.sqrt(2.0) - 1.0) * m * c * c, rtol=1e-15)  # p = mc


@pytest.mark.parametrize("T", [1, 2, 3, 5])
def test_repeated_splits_give_two_to_the_t_descendants(T):
    # E / e0 = 2^T * 1.01 with ratio 2: the target level is T for the parent and
    # every descendant (children keep the parent's baseline)
    e0 = 0.25
    target_e = e0 * 2.0 ** T * 1.01
    p = np.sqrt((target_e + 1.0) ** 2 - 1.0)
    n, cap = 3, 3 * 2 ** T
    parts = _set(n, cap + 1, p, e0, weight=[1.0, 0.5, 0.25])
    count, calls = n, 0
    while True:
        new, st = O.split_energy_growth(_props(), _policy(), parts, count, cap + 1, 1000)
        calls += 1
        assert st["active_count"] == count and st["skipped_capacity_count"] == 0
        if st["split_count"] == 0:
            break
        assert new == count + st["split_count"]
        count = new
    assert count == n * 2 ** T and calls == T + 1
    assert np.all(parts["split_level"][:count] == T)
    np.testing.assert_array_equal(np.sort(parts["weight"][:count]),
                                  np.sort(np.repeat([1.0, 0.5, 0.25], 2 ** T) / 2 ** T))
    assert parts["weight"][:count].sum() == 1.75
    assert np.all(parts["e0"][:count] == e0)
    # the first call's children got ids first_root, first_root + 1, ... in parent order
    np.testing.assert_array_equal(parts["id"][n:2 * n], [1000, 1001, 1002])


def test_split_counters_and_guards():
    ratio = 2.0
    e_grow = 0.5 * 3.0  # E / e0 = 3 -> target level 1
    p_grow = np.sqrt((e_grow + 1.0) ** 2 - 1.0)
    n, cap = 8, 11
    parts = _set(n, cap, p_grow, 0.5)
    parts["e0"][1] = 0.0          # invalid baseline -> set to E, counted
    parts["e0"][2] = np.nan       # invalid baseline
    parts["status"][3] = 2        # not alive
    parts["split_level"][4] = 1   # target 1 <= level 1: no split
    parts["weight"][5] = 0.1      # child 0.05 < minimum 0.06
    parts["p"][6] = 0.1           # E / e0 < ratio
    start = {k: v.copy() for k, v in parts.items()}
    new, st = O.split_energy_growth(_props(), _policy(ratio, 0.06), parts, n, cap, 50)
    # splitting parents: 0 and 7 -> two children at slots 8 and 9
    assert st == {"active_count": 7, "split_count": 2, "skipped_capacity_count": 0,
                  "skipped_invalid_baseline_count": 2, "skipped_minimum_weight_count": 1}
    assert new == 10
    e = _ek(p_grow)
    assert parts["e0"][1] == e and parts["e0"][2] == e
    for k in ("x", "p", "weight", "split_level"):
        assert parts[k][3] == start[k][3] and parts[k][4] == start[k][4] and parts[k][5] == start[k][5]
    assert parts["weight"][0] == parts["weight"][8] == 0.5 and parts["split_level"][8] == 1
    assert parts["id"][8] == 50 and parts["id"][9] == 51
    # the repaired baselines now equal E: no split on the next pass
    new2, st2 = O.split_energy_growth(_props(), _policy(ratio, 0.06), parts, new, cap, 60)
    assert new2 == new and st2["split_count"] == 0 and st2["skipped_invalid_baseline_count"] == 0


def test_split_capacity_leaves_late_parents_untouched():
    e = 0.5 * 3.0
    p = np.sqrt((e + 1.0) ** 2 - 1.0)
    parts = _set(4, 6, p, 0.5)
    new, st = O.split_energy_growth(_props(), _policy(), parts, 4, 6, 0)
    assert new == 6 and st["split_count"] == 2 and st["skipped_capacity_count"] == 2
    np.testing.assert_array_equal(parts["weight"], [0.5, 0.5, 1.0, 1.0, 0.5, 0.5])
    np.testing.assert_array_equal(parts["split_level"], [1, 1, 0, 0, 1, 1])


def test_no_split_when_full_or_empty():
    parts = _set(4, 4, 10.0, 0.01)
    n, st = O.split_energy_growth(_props(), _policy(), parts, 4, 4, 0)
    assert n == 4 and all(v == 0 for v in st.values())
    n, st = O.split_energy_growth(_props(), _policy(), parts, 0, 4, 0)
    assert n == 0 and st["active_count"] == 0


def test_boundaries_wrap_and_escape():
    xe, ye = np.linspace(-1.0, 1.0, 9), np.linspace(-0.5, 0.5, 5)
    x = np.array([-1.0, 1.0, 2.75, -3.2, 0.3, 0.3, np.nan, 0.0])
    y = np.array([0.0, 0.5, -0.5, 0.2, 0.5000001, -0.7, 0.0, 9.0])
    n = len(x)
    for mask in range(4):
        parts = _set(n, n, 1.0, 1.0)
        parts["x"][:], parts["y"][:] = x, y
        parts["status"][7] = 2
        esc = O.apply_boundaries(xe, ye, mask, parts, n)
        for d, (key, lo, hi, v) in enumerate((("x", -1.0, 1.0, x), ("y", -0.5, 0.5, y))):
            got = parts[key]
            alive = parts["status"] != 2
            if mask & (1 << d):
                ok = alive & np.isfinite(v)
                assert np.all((got[ok] >= lo) & (got[ok] < hi))
                k = (v[ok] - got[ok]) / (hi - lo)
                np.testing.assert_allclose(k, np.round(k), atol=1e-12)
        out = np.zeros(n, bool)
        for d, (lo, hi, v) in enumerate(((-1.0, 1.0, x), (-0.5, 0.5, y))):
            if not mask & (1 << d):
                out |= ~((v >= lo) & (v <= hi))
        out[7] = False
        assert esc == out.sum()
        np.testing.assert_array_equal(parts["status"], np.where(out, 1, np.where(np.arange(n) == 7, 2, 0)))
