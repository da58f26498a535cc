"""Stretched grids on the GPU (faces refined toward y = 0, max/min spacing 10, as
in the production Parker shape): the packed coefficients of both models, and
pushes with injected noise (Parker and focused, FP64, adaptive dt, time
blending) match the oracle to 1e-12; the FP32 push follows the FP64 one; the
compact reader rejects stretched faces unless asked, the generic solvers accept
them (their parity on stretched grids is in the generic and focus tests) and the
ensemble refuses them."""
import numpy as np
import pytest

import kpt
import kpt_oracle as O
import workloads as W
from helpers import SQRT3, copy_parts, random_particles

pytestmark = pytest.mark.gpu

REL = 1e-12


def _field(nx=48, ny=40, seed=7):
    ye = W.stretched_edges(ny)
    xe, ye, planes = W.island_field(nx, ny, width=0.05, seed=seed, y_edges=ye)
    planes[2] = planes[2] + 0.1
    _, _, upper = W.island_field(nx, ny, width=0.05, seed=seed + 1, y_edges=ye)
    upper[2] = upper[2] + 0.1
    return xe, ye, planes, upper


def _close(a, b, s, keys=("x", "y", "p", "t", "step_dt", "mu")):
    scale = {"x": s.x_width, "y": s.y_width, "mu": 1.0}
    for k in keys:
        den = np.maximum(np.maximum(np.abs(a[k]), np.abs(b[k])), scale.get(k, 0.0))
        den = np.where(den == 0, 1.0, den)
        err = np.abs(a[k] - b[k]) / den
        assert err.max() <= REL, f"{k}: max rel err {err.max():.3e} at {int(err.argmax())}"
    np.testing.assert_array_equal(a["status"], b["status"])


@pytest.mark.parametrize("model", [16, 21])
def test_stretched_pack_matches_oracle(model):
    xe, ye, planes, _ = _field()
    with kpt.Context(16) as c:
        c.field_upload(kpt.SLOT_LOWER, model, xe, ye, planes)
        got = c.field_coefficients(kpt.SLOT_LOWER, len(xe) - 1, len(ye) - 1, model)
    want = O.Field(xe, ye, planes, model).coefficients()
    scale = np.maximum(np.abs(want).max(axis=(0, 1)), 1e-300)
    assert (np.abs(got - want) / scale).max() < 1e-13


@pytest.mark.parametrize("model,focused,closure", [(16, False, None), (21, True, None), (21, True, "qlt")])
@pytest.mark.parametrize("steps", [1, 12])
def test_stretched_push_injected_noise_matches_oracle(model, focused, closure, steps):
    xe, ye, planes, upper = _field()
    h = kpt.default_host_settings()
    h.focused_transport = int(focused)
    h.p0, h.mhd_dt, h.dt_min_rel, h.dt_max_rel = 1.0, 1e-3, 1e-3, 1e-1
    h.duu0, h.particle_v0 = 5.0, 0.5
    h.max_particle_steps_per_interval = steps
    s = kpt.make_device_settings(h, xe, ye)
    assert s.interpolation_shift_y == 0.0
    if closure == "qlt":
        s.closure = kpt.CLOSURE_QLT
        s.qlt_sigma2, s.qlt_correlation_length = 0.1, 0.02
    rng = np.random.default_rng(21)
    n = 4096
    parts = random_particles(n, s, rng, p_range=(0.3, 3.0))
    parts["y"][:] = rng.uniform(-0.08, 0.08, n)  # where the faces are finest
    parts["t"][:] = rng.uniform(0, 0.5 * s.mhd_dt, n)
    if focused:
        parts["mu"][:] = rng.uniform(-0.95, 0.95, n)
    noise = rng.uniform(-SQRT3, SQRT3, size=(n, steps, 4))
    want = copy_parts(parts)
    lo, hi = O.Field(xe, ye, planes, model), O.Field(xe, ye, upper, model)
    o_steps, o_fail = O.push_interval(lo, hi, s, 0.0, 3, 1, want, noise=noise)
    with kpt.Context(n) as c:
        c.field_upload(kpt.SLOT_LOWER, model, xe, ye, planes)
        c.field_upload(kpt.SLOT_UPPER, model, xe, ye, upper)
        c.upload(parts)
        c.set_noise(noise)
        st = c.push_interval(s, 0.0, 3)
        got = c.download()
    assert st.particle_steps == o_steps and st.failures == o_fail
    _close(got, want, s)


def test_stretched_fp32_follows_fp64():
    xe, ye, planes, upper = _field(96, 80)
    h = W.parker_harris_host_settings()
    s = kpt.make_device_settings(h, xe, ye)
    n = 20000
    rng = np.random.default_rng(4)
    parts = random_particles(n, s, rng, p_range=(1.0, 1.0))
    parts["y"][:] = rng.uniform(-0.1, 0.1, n)
    out = {}
    for prec in (kpt.PREC_FP64, kpt.PREC_FP32):
        s2 = s.copy()
        s2.precision = prec
        s2.push_chunks = 4
        with kpt.Context(n, seed=99) as c:
            c.field_upload(kpt.SLOT_LOWER, 16, xe, ye, planes)
            c.field_upload(kpt.SLOT_UPPER, 16, xe, ye, upper)
            c.upload(parts)
            st = c.push_interval(s2, 0.0, 0)
            assert st.failures == 0
            g = c.download()
            order = np.argsort(g["id"])
            out[prec] = {k: v[order] for k, v in g.items()}
    a, b = out[kpt.PREC_FP64], out[kpt.PREC_FP32]
    # same Philox streams: FP32 rounding only, pathwise (1000 fixed steps)
    dx = np.abs(a["x"] - b["x"])
    dx = np.minimum(dx, s.x_width - dx)
    assert np.median(dx) < 1e-5 and np.median(np.abs(a["p"] / b["p"] - 1)) < 1e-5


def test_stretched_faces_accepted_or_refused(tmp_path):
    xe, ye, planes, _ = _field(16, 12)
    path = str(tmp_path / "field00000.bin")
    kpt.compact_field_write(path, xe, ye, planes)
    with kpt.Context(16) as c:
        with pytest.raises(kpt.KptError):
            c.field_load_compact(path, kpt.SLOT_LOWER, 16)            # the reference reader's rule
        c.field_load_compact(path, kpt.SLOT_LOWER, 16, allow_stretched=True)
        c.generic_build(kpt.generic_default_params(), kpt.SLOT_LOWER)         # the generic solvers take it
        c.focus_build(kpt.generic_focus_default_params(), kpt.SLOT_LOWER)
        with kpt.Ensemble(c, [8]) as e:
            with pytest.raises(kpt.KptError):
                e.field_upload(0, 16, xe, ye, planes)
