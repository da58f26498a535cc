"""LDS staging of the frame pair: the FP32 and FP64 pushes sample the stencil
records a wavefront stages in LDS (a node box around its lanes' stencils) or, for
lanes outside the box, the node-major pair in global memory. The staged records
are copies and the sample's arithmetic is the same, so a push with staging must
equal the push without it (KPT_PUSH_STAGE=0 at context creation) bit for bit, on
every particle field and counter. (The FP64 pushes' agreement with the oracle,
1e-12 with injected noise, is tested in test_gpu_parity.py and
test_gpu_stretched.py through the same staged path.)

Cases: the focused push with the exact slab-QLT closure and with the reduced
closure, and the Parker push (adaptive dt), in FP32 and FP64, on the
tanh-stretched and the uniform island field; cell-sorted chunks (compact
wavefronts: most lanes staged) and one unsorted chunk of randomly placed
particles (dispersed wavefronts: most lanes read global memory)."""
import os

import numpy as np
import pytest

import kpt
import workloads as W

pytestmark = pytest.mark.gpu

N = 32768


def _field(grid):
    if grid == "stretched":
        ye = W.stretched_edges(256, ratio=10.0)
        xe, ye, f0 = W.island_field(512, 256, seed=7, y_edges=ye)
        _, _, f1 = W.island_field(512, 256, seed=8, y_edges=ye)
    else:
        xe, ye, f0 = W.island_field(512, 256, seed=7)
        _, _, f1 = W.island_field(512, 256, seed=8)
    return xe, ye, f0, f1


def _push(field, model, s, start, stage):
    xe, ye, f0, f1 = field
    old = os.environ.get("KPT_PUSH_STAGE")
    os.environ["KPT_PUSH_STAGE"] = "1" if stage else "0"
    try:
        with kpt.Context(len(start["x"]), seed=21) as c:
            c.field_upload(kpt.SLOT_LOWER, model, xe, ye, f0)
            c.field_upload(kpt.SLOT_UPPER, model, xe, ye, f1)
            c.upload(start)
            st = c.push_interval(s, 0.0, 0)
            got = c.download()
    finally:
        if old is None:
            del os.environ["KPT_PUSH_STAGE"]
        else:
            os.environ["KPT_PUSH_STAGE"] = old
    return got, st


def _equal(field, model, s, start):
    a, sa = _push(field, model, s, start, True)
    b, sb = _push(field, model, s, start, False)
    assert sa.particle_steps == sb.particle_steps and sa.failures == sb.failures
    assert sa.particle_steps > 0
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    return sa


PRECISIONS = pytest.mark.parametrize("prec", [kpt.PREC_FP32, kpt.PREC_FP64], ids=["fp32", "fp64"])


@PRECISIONS
@pytest.mark.parametrize("grid", ["stretched", "uniform"])
@pytest.mark.parametrize("closure", [kpt.CLOSURE_QLT, kpt.CLOSURE_REDUCED], ids=["qlt", "reduced"])
@pytest.mark.parametrize("chunks", [8, 1], ids=["sorted", "unsorted"])
def test_focused_staged_equals_global(prec, grid, closure, chunks):
    field = _field(grid)
    xe, ye = field[0], field[1]
    s = W.focused_qlt_device_settings(xe, ye, closure=closure, precision=prec)
    steps = 200
    s.mhd_dt = steps * s.dt_min
    s.max_particle_steps_per_interval = steps + 2
    s.push_chunks = chunks
    start = W.focused_qlt_particles(N, s, seed=4)
    st = _equal(field, 21, s, start)
    assert st.failures == 0


@PRECISIONS
@pytest.mark.parametrize("grid", ["stretched", "uniform"])
@pytest.mark.parametrize("chunks", [8, 1], ids=["sorted", "unsorted"])
def test_parker_staged_equals_global(prec, grid, chunks):
    field = _field(grid)
    xe, ye = field[0], field[1]
    h = kpt.default_host_settings()
    h.split_particles = 0
    h.mhd_dt = 0.01
    s = kpt.make_device_settings(h, xe, ye)
    s.precision = prec
    s.push_chunks = chunks
    rng = np.random.default_rng(5)
    start = kpt.empty_particles(N)
    start["x"][:] = rng.uniform(s.x_lower, s.x_upper, N)
    start["y"][:] = rng.uniform(s.y_lower, s.y_upper, N)
    start["p"][:] = np.exp(rng.uniform(0.0, np.log(2.0), N))
    start["weight"][:] = 1.0
    start["id"][:] = np.arange(N, dtype=np.uint64)
    st = _equal(field, 16, s, start)
    assert st.failures == 0
