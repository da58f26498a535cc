"""The generic particle generator on the GPU (kpt_candidates_*, kpt_inject_*,
kpt_recycle_*; include/ParticleGenerator.hpp) against the oracle: candidate
cells equal, cumulative weights to 1e-13 (a device scan against a sequential
sum); with the device's cumulative weights every injected or recycled particle
equals the oracle's bit for bit (same Philox streams), on uniform and stretched
grids; rejection counts agree; a candidate set keeps its grid when the slot is
reloaded; the reference's error messages."""
import math

import numpy as np
import pytest

import kpt
import kpt_oracle as O
import workloads as W

pytestmark = pytest.mark.gpu

NX, NY = 24, 20
FIELDS = ("x", "y", "p", "mu", "weight", "t", "step_dt", "status", "split_level", "id", "e0")
SEED = 31


def _field(stretched, seed=5):
    ye = W.stretched_edges(NY, -0.5, 0.5, 8.0) if stretched else None
    xe, ye2, planes = W.island_field(NX, NY, width=0.08, seed=seed, y_edges=ye)
    planes[2] = planes[2] + np.random.default_rng(seed).normal(0.0, 0.3, planes[2].shape)
    return xe, ye2, planes


def _kin(status=0):
    return kpt.injection_kinematics(momentum=2.0, mu=-0.4, weight=0.25, status=status, time=0.125, step_dt=2e-4)


E0 = (math.sqrt(1.0 + 2.0 * 2.0) - 1.0) * 1.0 * 1.0 * 1.0  # kinetic energy of |p| = 2, m = c = 1


def _ctx(xe, ye, planes, n_pre=100, cap=20000):
    c = kpt.Context(cap, seed=SEED)
    c.field_upload(kpt.SLOT_LOWER, 16, xe, ye, planes)
    if n_pre:
        s = kpt.make_device_settings(kpt.default_host_settings(), xe, ye)
        assert c.inject(s, n_pre, 0.0, 0) == n_pre
    return c


def _assert_equal(got, want, n, first=0):
    for k in FIELDS:
        np.testing.assert_array_equal(got[k][first:n], want[k][first:n], err_msg=k)


@pytest.mark.parametrize("stretched", [False, True])
@pytest.mark.parametrize("pred", [(2, kpt.CMP_GREATER, 0.1), (0, kpt.CMP_ABS_LESS, 0.5), (4, kpt.CMP_LESS_EQUAL, 0.0)])
def test_field_candidates_match_oracle(stretched, pred):
    xe, ye, planes = _field(stretched)
    region = (-0.81, 0.63, -0.37, 0.41)
    with _ctx(xe, ye, planes, 0) as c, c.candidates_from_field(kpt.SLOT_LOWER, pred, region) as cs:
        cells, cum = cs.download()
        total = cs.total
    w_cells, w_cum, w_total = O.candidates_field(planes, xe, ye, pred, region)
    assert len(w_cells) > 5
    np.testing.assert_array_equal(cells, w_cells)
    np.testing.assert_allclose(cum, w_cum, rtol=1e-13)
    assert total == pytest.approx(w_total, rel=1e-13)


@pytest.mark.parametrize("stretched", [False, True])
def test_grid_candidates_match_oracle(stretched):
    xe, ye, planes = _field(stretched)
    with _ctx(xe, ye, planes, 0) as c, c.candidates_from_grid(kpt.SLOT_LOWER) as cs:
        cells, cum = cs.download()
        assert cs.count == NX * NY
        total = cs.total
    w_cum, w_total = O.candidates_grid(xe, ye)
    np.testing.assert_array_equal(cells, np.arange(NX * NY))
    np.testing.assert_allclose(cum, w_cum, rtol=1e-13)
    assert total == pytest.approx((xe[-1] - xe[0]) * (ye[-1] - ye[0]), rel=1e-13)


@pytest.mark.parametrize("stretched", [False, True])
@pytest.mark.parametrize("kind", ["field", "grid"])
def test_inject_from_candidates_bit_exact(stretched, kind):
    xe, ye, planes = _field(stretched)
    n = 6000
    with _ctx(xe, ye, planes) as c:
        cs = (c.candidates_from_field(kpt.SLOT_LOWER, (2, kpt.CMP_GREATER_EQUAL, 0.05)) if kind == "field"
              else c.candidates_from_grid(kpt.SLOT_LOWER))
        with cs:
            cells, cum = cs.download()
            call = c.generator_calls()
            assert c.inject_from_candidates(cs, n, _kin()) == n
            assert c.generator_calls() == call + 1 and c.count == 100 + n
            got = c.download()
            want = {k: v.copy() for k, v in got.items()}
            O.generator_inject_candidates(0 if kind == "field" else 1, cells if kind == "field" else None, cum,
                                          cs.total, xe, ye, n, _kin(), E0, SEED, call, 100, want, 100)
    _assert_equal(got, want, 100 + n)
    assert np.all(got["e0"][100:] == E0) and np.all(got["id"][100:] == np.arange(100, 100 + n))


@pytest.mark.parametrize("stretched", [False, True])
def test_inject_uniform_if_bit_exact(stretched):
    xe, ye, planes = _field(stretched)
    n, region, pred = 8000, (-0.9, 0.7, -0.45, 0.35), (2, kpt.CMP_GREATER, 0.2)
    with _ctx(xe, ye, planes) as c:
        call = c.generator_calls()
        injected, rejected = c.inject_uniform_if(kpt.SLOT_LOWER, n, _kin(), region=region, pred=pred, max_attempts=2)
        got = c.download()
    want = {k: v.copy() for k, v in got.items()}
    w_rej = O.generator_inject_uniform(planes, xe, ye, region, pred, 2, n, _kin(), E0, SEED, call, 100, want, 100)
    assert injected == n and rejected == w_rej and 0 < rejected < n
    _assert_equal(got, want, 100 + n)
    # no region, no predicate: the physical domain
    with _ctx(xe, ye, planes, 0) as c:
        assert c.inject_uniform_if(kpt.SLOT_LOWER, 2000, _kin()) == (2000, 0)
        got = c.download()
    want = {k: v.copy() for k, v in got.items()}
    assert O.generator_inject_uniform(planes, xe, ye, None, None, 64, 2000, _kin(), E0, SEED, 0, 0, want, 0) == 0
    _assert_equal(got, want, 2000)


@pytest.mark.parametrize("stretched", [False, True])
def test_recycle_bit_exact(stretched):
    xe, ye, planes = _field(stretched)
    n = 5000
    rng = np.random.default_rng(8)
    with _ctx(xe, ye, planes, 0, cap=n) as c:
        parts = kpt.empty_particles(n)
        parts["x"][:] = rng.uniform(-1, 1, n)
        parts["y"][:] = rng.uniform(-0.5, 0.5, n)
        parts["p"][:] = rng.uniform(0.5, 2.0, n)
        parts["status"][:] = rng.integers(0, 3, n)
        parts["id"][:] = rng.permutation(n).astype(np.uint64) + 7
        parts["split_level"][:] = rng.integers(0, 4, n)
        c.upload(parts)
        want = c.download()
        # uniform with a predicate over the escaped slots, then candidates over the rest
        call = c.generator_calls()
        rec, rej = c.recycle_uniform_if(kpt.SLOT_LOWER, _kin(status=0), kpt.RECYCLE_ESCAPED,
                                        pred=(2, kpt.CMP_LESS, -0.1), max_attempts=3)
        w_rec, w_rej = O.generator_recycle(1, _kin(status=0), E0, SEED, call, want, n, kpt.RECYCLE_ESCAPED, xe, ye,
                                           planes=planes, pred=(2, kpt.CMP_LESS, -0.1), max_attempts=3)
        assert (rec, rej) == (w_rec, w_rej) and rec > 0
        with c.candidates_from_field(kpt.SLOT_LOWER, (1, kpt.CMP_ABS_GREATER, 0.01)) as cs:
            cells, cum = cs.download()
            rec2 = c.recycle_from_candidates(cs, _kin(status=1), kpt.RECYCLE_ANY)
            w_rec2, _ = O.generator_recycle(0, _kin(status=1), E0, SEED, call + 1, want, n, kpt.RECYCLE_ANY, xe, ye,
                                            kind=0, cells=cells, cum=cum, total=cs.total)
        assert rec2 == w_rec2
        got = c.download()
    _assert_equal(got, want, n)
    np.testing.assert_array_equal(got["id"], parts["id"])


def test_candidate_set_keeps_its_grid_when_the_slot_is_reloaded():
    xe, ye, planes = _field(True)
    xu, yu, pu = _field(False, seed=9)
    with _ctx(xe, ye, planes, 0) as c, c.candidates_from_grid(kpt.SLOT_LOWER) as cs:
        cells, cum = cs.download()
        c.field_upload(kpt.SLOT_LOWER, 16, xu, yu, pu)   # same shape, uniform faces
        c.inject_from_candidates(cs, 3000, _kin())
        got = c.download()
    want = {k: v.copy() for k, v in got.items()}
    O.generator_inject_candidates(1, None, cum, cs.total, xe, ye, 3000, _kin(), E0, SEED, 0, 0, want, 0)
    _assert_equal(got, want, 3000)


def test_generator_errors():
    xe, ye, planes = _field(False)
    with _ctx(xe, ye, planes, 0, cap=100) as c:
        with c.candidates_from_field(kpt.SLOT_LOWER, (2, kpt.CMP_GREATER, 1e9)) as empty:
            assert empty.count == 0 and empty.total == 0.0
            with pytest.raises(kpt.KptError, match="field-conditioned injection has no candidate grid points"):
                c.inject_from_candidates(empty, 1, _kin())
            with pytest.raises(kpt.KptError, match="field-conditioned recycling has no candidate grid points"):
                c.recycle_from_candidates(empty, _kin(), kpt.RECYCLE_ANY)
        with c.candidates_from_grid(kpt.SLOT_LOWER) as g:
            with pytest.raises(kpt.KptError, match="requested range exceeds capacity"):
                c.inject_from_candidates(g, 101, _kin())
        bad = _kin()
        bad.weight = 0.0
        with pytest.raises(kpt.KptError, match="mu must be finite and weight must be positive"):
            c.inject_uniform_if(kpt.SLOT_LOWER, 1, bad)
        bad = _kin()
        bad.momentum = -1.0
        with pytest.raises(kpt.KptError, match="momentum magnitude must be finite and non-negative"):
            c.inject_uniform_if(kpt.SLOT_LOWER, 1, bad)
        with pytest.raises(kpt.KptError, match="bounds must be finite and ascending"):
            c.inject_uniform_if(kpt.SLOT_LOWER, 1, _kin(), region=(0.5, 0.1, 0.0, 0.1))
        with pytest.raises(kpt.KptError, match="component index is out of range"):
            c.candidates_from_field(kpt.SLOT_LOWER, (6, kpt.CMP_GREATER, 0.0))
        with pytest.raises(kpt.KptError, match="max_attempts must be positive"):
            c.inject_uniform_if(kpt.SLOT_LOWER, 1, _kin(), max_attempts=0)
        with pytest.raises(kpt.KptError, match="field slot is empty"):
            c.candidates_from_grid(kpt.SLOT_UPPER)
        assert c.count == 0 and c.generator_calls() == 0
