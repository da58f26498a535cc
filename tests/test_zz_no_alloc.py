"""kpt_push_interval_noalloc (the interval push with its launch arguments and
counters in the library's file-scope device array, no device allocation in the
call) against kpt_push_interval on the same inputs.

Two no-allocation calls in a row on one context, the second with fewer
particles, each against kpt_push_interval on a context of its own:
1. FP64 focused transport with the slab-QLT closure, time-blended frames, four
   push chunks, Philox noise (kpt_push_interval re-sorts by cell between chunks,
   the no-allocation push keeps the order): every particle equal bit for bit by
   id, the stats equal.
2. FP32 focused transport, one frame, one chunk, fp64_null_b above every |B| so
   that every sub-step runs the FP64 near-null path, which reads the launch
   arguments through their device record: bit for bit, in slot order.
Then an FP32 time-blended push (the no-allocation push samples both frames
directly instead of the paired copy): within FP32 agreement of the paths;
push_chunks beyond the array's chunk records is refused."""
import numpy as np
import pytest

import accuracy as A
import kpt
import workloads as W

pytestmark = pytest.mark.gpu

NX, NY = 256, 128
FIELDS = ("x", "y", "p", "t", "step_dt", "weight", "mu", "status", "split_level", "id")


@pytest.fixture(scope="module")
def frames():
    xe, ye, lower = W.island_field(NX, NY, seed=3)
    _, _, upper = W.island_field(NX, NY, seed=3, frame=1)
    return xe, ye, lower, upper


def _settings(xe, ye, precision, blend, chunks, null_b=None):
    s = W.focused_qlt_device_settings(xe, ye, precision=precision)
    s.mhd_dt = 100 * s.dt_min  # about 100 fixed sub-steps
    s.time_interpolation = int(blend)
    s.push_chunks = chunks
    if null_b is not None:
        s.fp64_null_b = null_b
    return s


def _context(frames, n, seed=77):
    xe, ye, lower, upper = frames
    c = kpt.Context(n, seed=seed)
    c.field_upload(kpt.SLOT_LOWER, kpt.MODEL_FOCUSED21, xe, ye, lower)
    c.field_upload(kpt.SLOT_UPPER, kpt.MODEL_FOCUSED21, xe, ye, upper)
    return c


def _sub_steps(s):
    """Sub-steps of one particle from t = 0 over [0, mhd_dt) at the fixed dt, the
    last one truncated to the interval end (the push's arithmetic, in double)."""
    t, n = 0.0, 0
    while t < s.mhd_dt:
        t += s.dt_min if s.dt_min <= s.mhd_dt - t else s.mhd_dt - t
        n += 1
    return n


def _stats(st):
    return (st.particle_steps, st.failures, st.active_in, st.launches)


def _reference(frames, s, parts):
    with _context(frames, len(parts["x"])) as c:
        c.upload(parts)
        st = c.push_interval(s, 0.0, 5)
        return _stats(st), c.download()


def _same(got, want, by_id):
    if by_id:
        got, want = A.sort_by_id(got), A.sort_by_id(want)
    for k in FIELDS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_noalloc_push_matches_push_interval_on_two_calls(frames):
    xe, ye, _, _ = frames
    s1 = _settings(xe, ye, kpt.PREC_FP64, blend=True, chunks=4)
    p1 = W.focused_qlt_particles(200_000, s1, seed=5)
    s2 = _settings(xe, ye, kpt.PREC_FP32, blend=False, chunks=1, null_b=1.0e30)
    p2 = W.focused_qlt_particles(30_000, s2, seed=6)
    want1 = _reference(frames, s1, p1)
    want2 = _reference(frames, s2, p2)
    assert want1[0][0] == 200_000 * _sub_steps(s1) and want2[0][0] == 30_000 * _sub_steps(s2)
    with _context(frames, 200_000) as c:
        c.upload(p1)
        st1 = c.push_interval_noalloc(s1, 0.0, 5)
        got1 = c.download()
        c.upload(p2)  # the second call: fewer particles, one chunk record of the four used before
        st2 = c.push_interval_noalloc(s2, 0.0, 5)
        got2 = c.download()
    assert _stats(st1) == want1[0] and _stats(st2) == want2[0]
    _same(got1, want1[1], by_id=True)
    _same(got2, want2[1], by_id=False)


def test_noalloc_push_without_stats_and_fp32_blended(frames):
    xe, ye, _, _ = frames
    s = _settings(xe, ye, kpt.PREC_FP32, blend=True, chunks=2)
    parts = W.focused_qlt_particles(50_000, s, seed=8)
    want_stats, want = _reference(frames, s, parts)
    with _context(frames, 50_000) as c:
        c.upload(parts)
        assert c.push_interval_noalloc(s, 0.0, 5, stats=False) is None
        got = c.download()
        c.upload(parts)
        st = c.push_interval_noalloc(s, 0.0, 5)
        again = c.download()
    _same(again, got, by_id=False)  # with and without a stats record: the same push
    assert _stats(st) == want_stats  # fixed dt: the same sub-steps
    err = A.pathwise_medians(A.sort_by_id(got), A.sort_by_id(want), s.x_width, s.y_width)
    assert err["x"] < 1e-5 * s.x_width and err["y"] < 1e-5 * s.y_width and err["p"] < 1e-5, err
    assert err["mu"] < 1e-3, err


def test_noalloc_push_refuses_more_chunks_than_records(frames):
    xe, ye, _, _ = frames
    s = _settings(xe, ye, kpt.PREC_FP64, blend=False, chunks=kpt.NOALLOC_MAX_CHUNKS + 1)
    parts = W.focused_qlt_particles(1000, s, seed=9)
    with _context(frames, 1000) as c:
        c.upload(parts)
        with pytest.raises(kpt.KptError, match="exceeds the 16 chunk records"):
            c.push_interval_noalloc(s, 0.0, 5)
        s.push_chunks = kpt.NOALLOC_MAX_CHUNKS  # the maximum itself runs
        st = c.push_interval_noalloc(s, 0.0, 5)
    assert st.launches == kpt.NOALLOC_MAX_CHUNKS and st.particle_steps == 1000 * _sub_steps(s)
