"""Field streaming (kpt_field_stage_compact / kpt_field_advance_staged): a frame
staged in the background while a push runs gives the same packed coefficients
and the same push results as the synchronous load after kpt_field_advance, on
uniform and stretched grids; read errors surface at the advance and leave the
slots as they were; the call order is enforced."""
import os

import numpy as np
import pytest

import kpt
import workloads as W
from helpers import random_particles

pytestmark = pytest.mark.gpu


def _frames(d, n, stretched):
    ye = W.stretched_edges(40) if stretched else None
    paths = []
    for f in range(n):
        xe, ye2, planes = W.island_field(64, 40, width=0.05, seed=11 + f, y_edges=ye)
        planes[2] = planes[2] + 0.1
        path = os.path.join(d, f"field{f:05d}.bin")
        kpt.compact_field_write(path, xe, ye2, planes)
        paths.append(path)
    return paths, xe, ye2


@pytest.mark.parametrize("stretched", [False, True])
@pytest.mark.parametrize("model", [16, 21])
def test_staged_frames_equal_synchronous_loads(tmp_path, stretched, model):
    paths, xe, ye = _frames(str(tmp_path), 5, stretched)
    h = kpt.default_host_settings()
    h.focused_transport = int(model == 21)
    h.dt_min_rel = h.dt_max_rel = 1e-2  # 100 fixed sub-steps per interval
    s = kpt.make_device_settings(h, xe, ye)
    s.push_chunks = 2
    rng = np.random.default_rng(3)
    parts = random_particles(4000, s, rng, p_range=(0.5, 2.0))
    results = []
    for staged in (False, True):
        with kpt.Context(4000, seed=5) as c:
            c.field_load_compact(paths[0], kpt.SLOT_LOWER, model, allow_stretched=stretched)
            c.field_load_compact(paths[1], kpt.SLOT_UPPER, model, allow_stretched=stretched)
            c.upload(parts)
            coeffs = []
            if staged:
                c.field_stage_compact(paths[2], model, allow_stretched=stretched)
            for frame in range(3):
                c.push_interval(s, frame * s.mhd_dt, frame)   # overlaps the stage
                if staged:
                    t = c.field_advance_staged()
                    assert t >= 0.0
                    if frame + 3 < len(paths):
                        c.field_stage_compact(paths[frame + 3], model, allow_stretched=stretched)
                else:
                    c.field_advance()
                    c.field_load_compact(paths[frame + 2], kpt.SLOT_UPPER, model, allow_stretched=stretched)
                coeffs.append([c.field_coefficients(sl, 64, 40, model) for sl in (kpt.SLOT_LOWER, kpt.SLOT_UPPER)])
            results.append((c.download(), coeffs))
    (a, ca), (b, cb) = results
    for x, y in zip(ca, cb):
        np.testing.assert_array_equal(x[0], y[0])
        np.testing.assert_array_equal(x[1], y[1])
    for k in ("x", "y", "p", "t", "mu", "status", "id"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_stage_errors_and_call_order(tmp_path):
    (tmp_path / "s").mkdir()
    (tmp_path / "u").mkdir()
    paths, xe, ye = _frames(str(tmp_path / "s"), 2, True)
    uni, _, _ = _frames(str(tmp_path / "u"), 2, False)
    with kpt.Context(16) as c:
        c.field_load_compact(uni[0], kpt.SLOT_LOWER, 16)
        c.field_load_compact(uni[1], kpt.SLOT_UPPER, 16)
        before = [c.field_coefficients(sl, 64, 40, 16) for sl in (kpt.SLOT_LOWER, kpt.SLOT_UPPER)]
        with pytest.raises(kpt.KptError, match="no field stage"):
            c.field_advance_staged()
        # stretched faces without the flag: the reader's error, at the advance
        c.field_stage_compact(paths[0], 16)
        with pytest.raises(kpt.KptError, match="in flight"):
            c.field_stage_compact(paths[1], 16)   # one stage at a time
        with pytest.raises(kpt.KptError, match="uniformly spaced"):
            c.field_advance_staged()
        c.field_stage_compact(str(tmp_path / "missing.bin"), 16)
        with pytest.raises(kpt.KptError, match="failed to open"):
            c.field_advance_staged()
        after = [c.field_coefficients(sl, 64, 40, 16) for sl in (kpt.SLOT_LOWER, kpt.SLOT_UPPER)]
        for x, y in zip(before, after):
            np.testing.assert_array_equal(x, y)
        # a stage left in flight is joined by the context's destruction
        c.field_stage_compact(uni[0], 16)
