"""The focused pushes at full size, through properties that do not depend on the
particle count (the oracle's statistical comparison runs on samples of the same
workload in test_gpu_stretched_headline.py).

Workload: focused transport with slab-QLT scattering on the 1024 x 512
island-chain field, y faces tanh-stretched toward the sheet (spacing ratio 10),
two time-blended frames, mu ~ U[-1, 1], p log-uniform on [1, 1e3], fixed dt 5e-5
over one 0.1 interval, 16 cell-sorted chunks:
  - 8 388 608 particles, FP32 with the learned SiLU closure (the bench's
    measured leg);
  - 4 194 304 particles, FP64 with the exact closure.

Checked for every particle: exactly the interval's sub-steps (their total equals
the count times the per-particle steps the interval arithmetic gives), status 0,
t at the interval end (to rounding), |mu| <= 1, p above the floor, weights and
ids unchanged (ids a permutation of the start ids after the cell sorts),
positions inside the periodic domain; and the momentum histogram's total equal
to the in-range weight."""
import numpy as np
import pytest

import kpt
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def field():
    ye = W.stretched_edges(512, ratio=10.0)
    xe, ye, f0 = W.island_field(1024, 512, seed=7, y_edges=ye)
    _, _, f1 = W.island_field(1024, 512, seed=8, y_edges=ye)
    return xe, ye, f0, f1


def _steps_per_particle(s):
    """Sub-steps of one particle from t = 0 to the interval end at fixed dt, with
    the last step truncated to the end (the mover's loop in double)."""
    t, n = 0.0, 0
    end = s.mhd_dt
    while t < end:
        dt = min(s.dt_min, end - t)
        t += dt
        n += 1
    return n


def _check(s, start, got, st, n):
    per = _steps_per_particle(s)
    assert st.failures == 0 and st.active_in == n
    assert st.particle_steps == n * per
    assert np.all(got["status"] == 0)
    # the last step is truncated to the interval end: t lands on it to rounding
    assert np.all(got["t"] >= s.mhd_dt) and np.all(got["t"] - s.mhd_dt <= 4e-16 * s.mhd_dt)
    assert np.all(np.abs(got["mu"]) <= 1.0)
    assert np.all(got["p"] >= 0.25 * s.p0) and np.all(np.isfinite(got["p"]))
    assert np.all((got["x"] >= s.x_lower) & (got["x"] <= s.x_upper))
    assert np.all((got["y"] >= s.y_lower) & (got["y"] <= s.y_upper))
    order_got, order_start = np.argsort(got["id"]), np.argsort(start["id"])
    np.testing.assert_array_equal(got["id"][order_got], start["id"][order_start])
    np.testing.assert_array_equal(got["weight"][order_got], start["weight"][order_start])


def _run(field, n, s, net=None):
    xe, ye, f0, f1 = field
    start = W.focused_qlt_particles(n, s, seed=11)
    with kpt.Context(n, seed=114514) as c:
        if net is not None:
            c.mlp_set(*net)
        c.field_upload(kpt.SLOT_LOWER, kpt.MODEL_FOCUSED21, xe, ye, f0)
        c.field_upload(kpt.SLOT_UPPER, kpt.MODEL_FOCUSED21, xe, ye, f1)
        c.upload(start)
        st = c.push_interval(s, 0.0, 0)
        got = c.download()
        pmin, pmax = 0.5, 2.0e3
        h = c.histogram_logp(128, pmin, pmax)
    inside = (got["p"] >= pmin) & (got["p"] <= pmax) & (got["status"] == 0)
    assert h.sum() == pytest.approx(got["weight"][inside].sum(), rel=1e-12)
    return start, got, st


def test_learned_closure_push_8m(field):
    xe, ye = field[0], field[1]
    o = kpt.mlp_default_train_opts()
    o.steps = 1000  # accuracy is not under test here
    with kpt.Context(16) as c:
        c.mlp_train(o)
        net = c.mlp_get()
    s = W.focused_qlt_device_settings(xe, ye, closure=kpt.CLOSURE_MLP, precision=kpt.PREC_FP32)
    s.push_chunks = 16
    n = W.FOCUSED_QLT_1024X512_8M["particles"]
    start, got, st = _run(field, n, s, net)
    _check(s, start, got, st, n)


def test_fp64_exact_push_4m(field):
    xe, ye = field[0], field[1]
    s = W.focused_qlt_device_settings(xe, ye, closure=kpt.CLOSURE_QLT, precision=kpt.PREC_FP64)
    s.push_chunks = 16
    n = W.FOCUSED_QLT_1024X512["particles"]
    start, got, st = _run(field, n, s)
    _check(s, start, got, st, n)


@pytest.mark.parametrize("prec", [kpt.PREC_FP64, kpt.PREC_FP32], ids=["fp64", "fp32"])
def test_parker_push_4m(field, prec):
    """The Parker push on 4 194 304 particles injected over the domain at the
    interval start (the bench's push-alone workload), adaptive dt: no failures,
    every particle at the interval end with at least one sub-step and at most
    the step guard, p above the floor, positions in the domain, ids and
    weights preserved."""
    xe, ye, f0, f1 = field
    h = kpt.default_host_settings()
    h.split_particles = 0
    s = kpt.make_device_settings(h, xe, ye)
    s.precision = prec
    s.push_chunks = 8
    n = 4 * 1024 * 1024
    with kpt.Context(n, seed=114514) as c:
        c.field_upload(kpt.SLOT_LOWER, kpt.MODEL_PARKER16, xe, ye, f0)
        c.field_upload(kpt.SLOT_UPPER, kpt.MODEL_PARKER16, xe, ye, f1)
        o = kpt.InjectOpts()
        o.fixed_time = 1
        assert c.inject(s, n, 0.0, 0, o) == n
        start = c.download()
        st = c.push_interval(s, 0.0, 0)
        got = c.download()
    assert st.failures == 0 and st.active_in == n
    assert n <= st.particle_steps <= n * s.max_particle_steps_per_interval
    assert np.all(got["status"] == 0)
    assert np.all(got["t"] >= s.mhd_dt) and np.all(got["t"] - s.mhd_dt <= 4e-16 * s.mhd_dt)
    assert np.all(got["p"] >= 0.25 * s.p0) and np.all(np.isfinite(got["p"]))
    assert np.all((got["x"] >= s.x_lower) & (got["x"] <= s.x_upper))
    assert np.all((got["y"] >= s.y_lower) & (got["y"] <= s.y_upper))
    order_got, order_start = np.argsort(got["id"]), np.argsort(start["id"])
    np.testing.assert_array_equal(got["id"][order_got], start["id"][order_start])
    np.testing.assert_array_equal(got["weight"][order_got], start["weight"][order_start])
