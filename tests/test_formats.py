"""Host-side formats of the push boundary: compact-field input, v4 snapshots,
histogram and summary CSVs. Pinned against the byte sizes and the CSV files of the
reference's tracked run (kokkos_cpu_format_example_frame180_262k/, copied under
tests/golden/ as data). No GPU is needed."""
import os

import numpy as np
import pytest

import kpt
from helpers import random_particles, uniform_edges

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_compact_field_roundtrip_and_size(kpt_mod, tmp_path):
    rng = np.random.default_rng(0)
    nx, ny = 37, 21
    xe, ye = uniform_edges(nx, -1, 1), uniform_edges(ny, 0, 0.7)
    planes = [rng.standard_normal((ny, nx)) for _ in range(6)]
    path = str(tmp_path / "field00000.bin")
    kpt.compact_field_write(path, xe, ye, planes)
    assert os.path.getsize(path) == 8 + 8 * (nx + ny + 2) + 48 * nx * ny
    xe2, ye2, p2 = kpt.compact_field_read(path)
    np.testing.assert_array_equal(xe2, xe)
    np.testing.assert_array_equal(ye2, ye)
    for a, b in zip(planes, p2):
        np.testing.assert_array_equal(a, b)


def test_compact_field_size_of_reference_frame(kpt_mod, tmp_path):
    """1024 x 1024 frame: 50 348 056 bytes (format_summary.txt:4-14)."""
    n = 1024
    e = uniform_edges(n, 0.0, 2.0)
    z = np.zeros((n, n))
    path = str(tmp_path / "f.bin")
    kpt.compact_field_write(path, e, e, [z] * 6)
    assert os.path.getsize(path) == 50348056


def test_compact_field_validation_messages(kpt_mod, tmp_path):
    nx, ny = 8, 4
    xe, ye = uniform_edges(nx, 0, 1), uniform_edges(ny, 0, 1)
    planes = [np.ones((ny, nx)) for _ in range(6)]
    # non-uniform edges: rejected by default (include/AthenaFieldReader.hpp:186-189)
    xs = xe.copy()
    xs[3] += 0.01
    path = str(tmp_path / "a.bin")
    kpt.compact_field_write(path, xs, ye, planes)
    with pytest.raises(kpt.KptError, match="uniformly spaced"):
        kpt.compact_field_read(path)
    kpt.compact_field_read(path, allow_nonuniform=True)
    # non-finite value
    bad = [p.copy() for p in planes]
    bad[4][1, 2] = np.nan
    kpt.compact_field_write(path, xe, ye, bad)
    with pytest.raises(kpt.KptError, match="non-finite Vy value"):
        kpt.compact_field_read(path)
    # trailing payload
    kpt.compact_field_write(path, xe, ye, planes)
    with open(path, "ab") as f:
        f.write(b"\0")
    with pytest.raises(kpt.KptError, match="trailing payload"):
        kpt.compact_field_read(path)
    # truncated
    kpt.compact_field_write(path, xe, ye, planes)
    with open(path, "r+b") as f:
        f.truncate(os.path.getsize(path) - 8)
    with pytest.raises(kpt.KptError, match="failed to read Vz"):
        kpt.compact_field_read(path)
    # descending edges
    kpt.compact_field_write(path, xe[::-1], ye, planes)
    with pytest.raises(kpt.KptError, match="strictly ascending"):
        kpt.compact_field_read(path)


def _settings():
    h = kpt.default_host_settings()
    return kpt.make_device_settings(h, uniform_edges(16, 0, 2), uniform_edges(16, 0, 2))


def test_snapshot_v4_layout(kpt_mod, tmp_path):
    s = _settings()
    n = 1000
    parts = random_particles(n, s, np.random.default_rng(1))
    parts["status"][::7] = 2
    parts["split_level"][:] = np.arange(n) % 5
    path = str(tmp_path / "particles_00003.bin")
    kpt.snapshot_write_host(path, parts, n, 4096, -1.0, 2.0, 0.1)
    raw = open(path, "rb").read()
    assert len(raw) == 88 + 100 * n
    assert raw[:8] == b"KPTPRT\0\0"
    hdr = np.frombuffer(raw[8:24], dtype=np.uint32)
    assert hdr[0] == 4 and hdr[1] == 0x01020304
    assert np.frombuffer(raw[16:24], dtype=np.int32).tolist() == [2, 3]
    assert np.frombuffer(raw[24:40], dtype=np.uint64).tolist() == [n, 4096]
    assert np.frombuffer(raw[40:88], dtype=np.float64).tolist() == [1.0, -1.0, 1.0, 1.0, 2.0, 0.0]
    rec = np.frombuffer(raw[88:], dtype=np.dtype([
        ("id", "<u8"), ("status", "<i4"), ("level", "<i4"), ("key", "<i4"),
        ("pos", "<f8", 2), ("prev", "<f8", 2), ("prev_step", "<f8", 2),
        ("p", "<f8"), ("mu", "<f8"), ("w", "<f8"), ("e0", "<f8")]))
    assert rec.itemsize == 100
    np.testing.assert_array_equal(rec["id"], np.arange(n))
    np.testing.assert_array_equal(rec["status"], parts["status"])
    np.testing.assert_array_equal(rec["level"], parts["split_level"])
    assert np.all(rec["key"] == 0)
    np.testing.assert_array_equal(rec["pos"][:, 0], parts["x"])
    np.testing.assert_array_equal(rec["prev"][:, 1], parts["y"])
    np.testing.assert_array_equal(rec["prev_step"][:, 0], parts["x"])
    np.testing.assert_array_equal(rec["p"], parts["p"])
    np.testing.assert_array_equal(rec["mu"], parts["mu"])
    np.testing.assert_array_equal(rec["w"], parts["weight"])
    assert np.all(rec["e0"] == np.sqrt(1 + 0.1 ** 2) - 1)
    back = kpt.snapshot_read_host(path, -1.0, 2.0)
    for k in ("x", "y", "p", "mu", "weight", "status", "split_level"):
        np.testing.assert_array_equal(back[k], parts[k])


def test_snapshot_size_of_reference_run(kpt_mod, tmp_path):
    """262 144 particles: 26 214 488 bytes (format_summary.txt:16-34)."""
    n = 262144
    parts = kpt.empty_particles(n)
    path = str(tmp_path / "particles_00181.bin")
    kpt.snapshot_write_host(path, parts, n, 1048576, -1.0, 2.0, 0.1)
    assert os.path.getsize(path) == 26214488


def test_snapshot_restart_checks(kpt_mod, tmp_path):
    s = _settings()
    parts = random_particles(10, s, np.random.default_rng(0))
    path = str(tmp_path / "particles_00001.bin")
    kpt.snapshot_write_host(path, parts, 10, 10, -1.0, 2.0, 0.1)
    with pytest.raises(kpt.KptError, match="charge does not match"):
        kpt.snapshot_read_host(path, 1.0, 2.0)
    with pytest.raises(kpt.KptError, match="split ratio does not match"):
        kpt.snapshot_read_host(path, -1.0, 1.5)
    with open(path, "r+b") as f:
        f.seek(8)
        f.write(np.uint32(3).tobytes())
    with pytest.raises(kpt.KptError, match="version 4"):
        kpt.snapshot_read_host(path, -1.0, 2.0)
    with open(path, "r+b") as f:
        f.write(b"XXXX")
    with pytest.raises(kpt.KptError, match="not a repository KPTPRT"):
        kpt.snapshot_read_host(path, -1.0, 2.0)


def test_histogram_csv_matches_reference_bytes(kpt_mod, tmp_path):
    """The 128-bin CSV of the tracked run (momentum_histogram_00181.csv), rebuilt
    from its own weights with pmin = 0.01, pmax = 10 (src/main.cpp:138-139, :1910)."""
    golden = os.path.join(GOLDEN, "ref_momentum_histogram_00181.csv")
    rows = open(golden).read().splitlines()[1:]
    weights = np.array([float(r.split(",")[3]) for r in rows])
    assert len(weights) == 128 and weights[42] == 261875
    out = str(tmp_path / "momentum_histogram_00181.csv")
    kpt.write_histogram_csv(out, 0.01, 10.0, weights)
    assert open(out, "rb").read() == open(golden, "rb").read()


def test_summary_csv_matches_reference(kpt_mod):
    golden = open(os.path.join(GOLDEN, "ref_summary_frame181.csv")).read().splitlines()
    assert kpt.summary_csv_header() == golden[0] + "\n"
    f = golden[1].split(",")
    sm = kpt.Summary()
    sm.stored, sm.active, sm.inactive = int(f[2]), int(f[3]), int(f[4])
    (sm.weight_sum, sm.p_min, sm.p_max, sm.p_avg, sm.dt_min, sm.dt_max,
     sm.dt_avg) = map(float, f[5:12])
    sm.max_split_level = int(f[12])
    tm = kpt.FrameTiming()
    tm.injected, tm.split, tm.skipped_split_capacity = int(f[13]), int(f[14]), int(f[15])
    tm.read_seconds, tm.inject_seconds, tm.move_seconds, tm.split_seconds = map(float, f[16:20])
    row = kpt.summary_csv_row(int(f[0]), float(f[1]), sm, tm, float(f[20]), float(f[21]))
    assert row == golden[1] + "\n"


def test_make_device_settings_matches_driver(kpt_mod):
    """make_device_settings (src/main.cpp:595-636) on the tracked run's grid
    (1024 x 1024 over [0, 2]^2) with the default host settings."""
    h = kpt.default_host_settings()
    e = uniform_edges(1024, 0.0, 2.0)
    s = kpt.make_device_settings(h, e, e)
    assert (s.x_lower, s.x_upper, s.y_lower, s.y_upper) == (0.0, 2.0, 0.0, 2.0)
    assert s.dx == 2.0 / 1024 and s.interpolation_shift_x == 0.5 * s.dx
    assert s.pindex == 3.0 - 1.6666667 and s.dt_min == 1e-7 * 0.1 and s.dt_max == 1e-2 * 0.1
    assert s.pmin_split == 2.0 * 0.1 and s.mu_max == 0.99
    assert s.pitch_angle_scattering_h0 == 0.2 and s.max_particle_steps_per_interval == 200000
    assert s.kpara0 == 0.00743592 and s.kperp_over_kpara == 0.01 and s.charge == -1.0


def test_make_fields_writes_a_frame_series(kpt_mod, tmp_path):
    """make_fields.py: compact files the driver's reader accepts (stretched ones
    only with the non-uniform flag), frames that differ by the drifting island
    phases, and links for frames past --distinct."""
    import make_fields
    import workloads as W

    out = tmp_path / "series"
    make_fields.main([str(out), "--frames", "5", "--nx", "32", "--ny", "16", "--distinct", "3"])
    names = sorted(p.name for p in out.iterdir())
    assert names == [f"field{f:05d}.bin" for f in range(5)]
    assert os.path.islink(out / "field00003.bin") and os.readlink(out / "field00003.bin") == "field00000.bin"
    assert os.path.getsize(out / "field00001.bin") == 8 + 8 * (32 + 16 + 2) + 48 * 32 * 16
    with pytest.raises(kpt.KptError, match="uniformly spaced"):
        kpt.compact_field_read(str(out / "field00000.bin"))
    f0 = kpt.compact_field_read(str(out / "field00000.bin"), allow_nonuniform=True)
    f1 = kpt.compact_field_read(str(out / "field00001.bin"), allow_nonuniform=True)
    np.testing.assert_array_equal(f0[1], W.stretched_edges(16))
    np.testing.assert_array_equal(f0[2][0], f1[2][0])       # Bx = tanh(y/w) does not drift
    assert not np.array_equal(f0[2][1], f1[2][1])            # the island chain does
    xe, ye, planes = W.island_field(32, 16, y_edges=W.stretched_edges(16), frame=0)
    np.testing.assert_array_equal(f0[2][1], planes[1])
    make_fields.main([str(tmp_path / "u"), "--frames", "1", "--nx", "8", "--ny", "4", "--grid", "uniform"])
    kpt.compact_field_read(str(tmp_path / "u" / "field00000.bin"))
