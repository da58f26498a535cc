"""Edge cases of the push and the per-particle kernels: an empty particle set
through every precision and closure (push, allocation-free push, split, sort,
histogram, summary), and ragged counts that leave the last wavefront partly
empty (1, 63, 65 and 129 particles). The FP64 pushes of ragged counts match the
oracle to 1e-12 with injected noise; the FP32 pushes, whose loop runs
wave-uniformly when the stencil records are staged in LDS (lanes past the count
take part in the staging but not in the push), equal the unstaged push bit for
bit; the learned-closure push, whose MFMAs take dummy features from empty
lanes, gives each particle of a ragged set exactly its result in full
wavefronts."""
import os

import numpy as np
import pytest

import kpt
import kpt_oracle as O
import workloads as W
from helpers import SQRT3, copy_parts, random_particles

pytestmark = pytest.mark.gpu

REL = 1e-12
RAGGED = [1, 63, 65, 129]


def _fields(model):
    xe, ye, planes = W.harris_field(64, 32)
    rng = np.random.default_rng(17)
    upper = [p + 0.02 * rng.standard_normal(p.shape) * (np.abs(p).max() + 0.1) for p in planes]
    return xe, ye, planes, upper


def _settings(xe, ye, focused, prec, closure=kpt.CLOSURE_REDUCED, steps=10):
    h = kpt.default_host_settings()
    h.focused_transport = int(focused)
    h.max_particle_steps_per_interval = steps
    h.p0 = 1.0
    h.mhd_dt = 1e-3
    h.dt_min_rel = 1e-3
    h.dt_max_rel = 1e-1
    h.duu0 = 5.0
    h.particle_v0 = 0.5
    s = kpt.make_device_settings(h, xe, ye)
    s.precision = prec
    if focused:
        s.closure = closure
        s.qlt_sigma2 = 0.1
        s.qlt_correlation_length = 0.02
    return s


CASES = [(16, False, kpt.CLOSURE_REDUCED), (21, True, kpt.CLOSURE_REDUCED), (21, True, kpt.CLOSURE_QLT)]


@pytest.mark.parametrize("prec", [kpt.PREC_FP64, kpt.PREC_FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("model,focused,closure", CASES, ids=["parker", "reduced", "qlt"])
def test_empty_particle_set(prec, model, focused, closure):
    xe, ye, planes, upper = _fields(model)
    s = _settings(xe, ye, focused, prec, closure)
    with kpt.Context(256, seed=3) as c:
        c.field_upload(kpt.SLOT_LOWER, model, xe, ye, planes)
        c.field_upload(kpt.SLOT_UPPER, model, xe, ye, upper)
        assert c.count == 0
        st = c.push_interval(s, 0.0, 0)
        assert (st.particle_steps, st.failures, st.active_in) == (0, 0, 0)
        st = c.push_interval_noalloc(s, 0.0, 0)
        assert (st.particle_steps, st.failures, st.active_in) == (0, 0, 0)
        c.split(s)
        assert c.count == 0
        assert c.sort(s) == 0
        assert c.sort(s, compact=True) == 0
        h = c.histogram_logp(16, 0.1, 10.0)
        assert np.all(np.asarray(h) == 0)
        sm = c.summary()
        assert sm["stored"] == 0 and sm["active"] == 0 and sm["weight_sum"] == 0.0
        got = c.download()
        assert all(len(v) == 0 for v in got.values())


def _ragged_case(n, model, focused, closure, prec, seed=23):
    xe, ye, planes, upper = _fields(model)
    s = _settings(xe, ye, focused, prec, closure)
    rng = np.random.default_rng(seed + n)
    parts = random_particles(n, s, rng, p_range=(0.3, 3.0))
    parts["t"][:] = rng.uniform(0, 0.5 * s.mhd_dt, n)
    return xe, ye, planes, upper, s, parts, rng


@pytest.mark.parametrize("n", RAGGED)
@pytest.mark.parametrize("model,focused,closure", CASES, ids=["parker", "reduced", "qlt"])
def test_ragged_fp64_matches_oracle(n, model, focused, closure):
    xe, ye, planes, upper, s, parts, rng = _ragged_case(n, model, focused, closure, kpt.PREC_FP64)
    steps = s.max_particle_steps_per_interval
    noise = rng.uniform(-SQRT3, SQRT3, size=(n, steps, 4))
    want = copy_parts(parts)
    lo, hi = O.Field(xe, ye, planes, model), O.Field(xe, ye, upper, model)
    o_steps, o_fail = O.push_interval(lo, hi, s, 0.0, 3, 1, want, noise=noise)
    with kpt.Context(n, seed=5) as c:
        c.field_upload(kpt.SLOT_LOWER, model, xe, ye, planes)
        c.field_upload(kpt.SLOT_UPPER, model, xe, ye, upper)
        c.upload(parts)
        c.set_noise(noise)
        st = c.push_interval(s, 0.0, 3)
        got = c.download()
    assert st.particle_steps == o_steps and st.failures == o_fail and st.active_in == n
    scale = {"x": s.x_width, "y": s.y_width, "mu": 1.0}
    for k in ("x", "y", "p", "t", "step_dt", "mu"):
        a, b = got[k], want[k]
        den = np.maximum(np.maximum(np.abs(a), np.abs(b)), scale.get(k, 0.0))
        den = np.where(den == 0, 1.0, den)
        err = np.abs(a - b) / den
        assert err.max() <= REL, f"{k}: max rel err {err.max():.3e}"
    np.testing.assert_array_equal(got["status"], want["status"])


def _fp32_push(xe, ye, planes, upper, model, s, parts, stage, capacity):
    old = os.environ.get("KPT_PUSH_STAGE")
    os.environ["KPT_PUSH_STAGE"] = "1" if stage else "0"
    try:
        with kpt.Context(capacity, seed=9) as c:
            c.field_upload(kpt.SLOT_LOWER, model, xe, ye, planes)
            c.field_upload(kpt.SLOT_UPPER, model, xe, ye, upper)
            c.upload(parts)
            st = c.push_interval(s, 0.0, 0)
            got = c.download()
    finally:
        if old is None:
            del os.environ["KPT_PUSH_STAGE"]
        else:
            os.environ["KPT_PUSH_STAGE"] = old
    return st, got


@pytest.mark.parametrize("n", RAGGED)
@pytest.mark.parametrize("model,focused,closure", CASES, ids=["parker", "reduced", "qlt"])
def test_ragged_fp32_staged_equals_unstaged(n, model, focused, closure):
    xe, ye, planes, upper, s, parts, _ = _ragged_case(n, model, focused, closure, kpt.PREC_FP32)
    s.max_particle_steps_per_interval = 200
    s.mhd_dt = 200 * s.dt_min
    parts["t"][:] = 0.0
    a_st, a = _fp32_push(xe, ye, planes, upper, model, s, parts, True, n + 64)
    b_st, b = _fp32_push(xe, ye, planes, upper, model, s, parts, False, n + 64)
    assert a_st.particle_steps == b_st.particle_steps > 0
    assert a_st.failures == b_st.failures == 0 and a_st.active_in == n
    assert len(a["x"]) == n
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


@pytest.fixture(scope="module")
def small_net():
    """A briefly trained closure network per activation (its accuracy is not
    under test here, only that a lane's result is independent of its wavefront)."""
    nets = {}
    for act in (kpt.MLP_SILU, kpt.MLP_HARD_SWISH):
        o = kpt.mlp_default_train_opts()
        o.samples = 1 << 16
        o.steps = 200
        o.hull.activation = act
        with kpt.Context(16) as c:
            c.mlp_train(o)
            nets[act] = c.mlp_get()
    return nets


@pytest.mark.parametrize("act", [kpt.MLP_SILU, kpt.MLP_HARD_SWISH], ids=["silu", "hard-swish"])
@pytest.mark.parametrize("grid", ["stretched", "uniform"])
@pytest.mark.parametrize("n", [1, 65])
def test_ragged_mlp_push_equals_full_wavefronts(small_net, act, grid, n):
    """The learned-closure push runs its sub-steps wave-uniformly (the MFMAs need
    all 64 lanes; lanes without a particle evaluate dummy features). Each
    particle's trajectory depends only on its own state and id, so a ragged set
    must give every particle exactly what it gets inside a set of full
    wavefronts (matched by id after the push's cell sorts)."""
    ye = W.stretched_edges(32, ratio=10.0) if grid == "stretched" else None
    xe, ye, f0 = W.island_field(64, 32, seed=3, y_edges=ye)
    _, _, f1 = W.island_field(64, 32, seed=4, y_edges=ye)
    s = W.focused_qlt_device_settings(xe, ye, closure=kpt.CLOSURE_MLP, precision=kpt.PREC_FP32)
    s.mhd_dt = 100 * s.dt_min
    full = W.focused_qlt_particles(192, s, seed=8)
    sub = {k: v[:n].copy() for k, v in full.items()}

    def run(parts):
        with kpt.Context(len(parts["x"]), seed=13) as c:
            c.mlp_set(*small_net[act])
            c.field_upload(kpt.SLOT_LOWER, 21, xe, ye, f0)
            c.field_upload(kpt.SLOT_UPPER, 21, xe, ye, f1)
            c.upload(parts)
            st = c.push_interval(s, 0.0, 0)
            return c.download(), st

    a, sa = run(sub)
    b, sb = run(full)
    assert sa.active_in == n and sb.active_in == 192 and sa.particle_steps > 0 and sa.failures == 0
    oa, ob = np.argsort(a["id"]), np.argsort(b["id"])
    for k in a:
        np.testing.assert_array_equal(a[k][oa], b[k][ob][:n], err_msg=k)
