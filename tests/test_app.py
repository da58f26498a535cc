"""The frame-loop driver kpt_run (run_transport_loop, src/main.cpp:2253-2429):
command-line validation on the CPU; on the GPU, a three-frame run on synthetic
compact-field files whose snapshot, histogram and summary rows equal a loop
driven through the library directly, and a restart that appends to the run."""
import os
import subprocess

import numpy as np
import pytest

import kpt
import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
APP = os.path.join(os.path.dirname(HERE), "kokkos-particletransport_amd", "kpt_run")


def _run(*args, timeout=120):
    return subprocess.run([APP, *args], capture_output=True, text=True, timeout=timeout)


@pytest.fixture(scope="module")
def app():
    if not os.path.exists(APP):
        pytest.fail(f"{APP} is missing: build it with `make -C kokkos-particletransport_amd`")
    return APP


def test_help_lists_reference_options(app):
    r = _run("--help")
    assert r.returncode == 0
    for opt in ("--profile", "--transport", "--restart-particle-snapshot", "--overwrite-output", "--frames",
                "--rank-scale", "--histogram-interval", "--particle-snapshot-interval", "--walltime-hours",
                "--walltime-reserve-minutes", "--no-split", "--no-time-interp", "--precision", "--closure",
                "--mlp-activation"):
        assert opt in r.stdout, opt


@pytest.mark.parametrize("args,msg", [
    (["--end-frame", "0"], "Reconnection2DSettings: end frame must exceed start frame."),
    (["--profile", "nope"], "Unknown run profile: nope."),
    (["--bogus"], "Unknown option: --bogus"),
    (["--frames"], "Missing value for --frames."),
    (["--capacity", "0"], "Reconnection2DSettings: particle counts and capacity must be positive."),
    (["--walltime-reserve-seconds", "5"], "reserve requires a positive walltime limit"),
    (["--split-ratio", "1.0"], "Reconnection2DSettings: invalid physical parameters."),
    (["--restart-particle-snapshot", "/nonexistent/particles_00002.bin", "--frames", "1"],
     "Reconnection2DSettings: restart particle snapshot does not exist."),
    (["--restart-particle-snapshot", "/tmp/snap.bin"],
     "Cannot infer restart frame from particle snapshot name; expected particles_XXXXX.bin."),
    (["--closure", "mlp", "--transport", "focused"], "the learned pitch-angle closure runs with --precision fp32"),
    (["--mlp-activation", "tanh"], "Unknown network activation: tanh"),
])
def test_validation_messages(app, args, msg):
    r = _run(*args)
    assert r.returncode == 1
    assert msg in r.stderr, r.stderr


def _write_fields(d, frames, nx=64, ny=32):
    for f in range(frames):
        xe, ye, planes = W.harris_field(nx, ny, inflow=0.1 + 0.01 * f)
        kpt.compact_field_write(os.path.join(d, f"field{f:05d}.bin"), xe, ye, planes)


@pytest.mark.gpu
def test_app_matches_library_loop_and_restarts(app, tmp_path):
    fields = tmp_path / "fields"
    out = tmp_path / "out"
    fields.mkdir()
    _write_fields(str(fields), 5)
    ppf, cap, seed = 2048, 16384, 7
    common = ["--profile", "smoke", "--field-dir", str(fields), "--output-dir", str(out),
              "--particles-per-frame", str(ppf), "--capacity", str(cap), "--seed", str(seed)]
    r = _run(*common, "--frames", "3")
    assert r.returncode == 0, r.stderr
    assert "Finished in" in r.stdout
    for name in ("summary.csv", "run.log", "momentum_histogram_00003.csv", "particles_00003.bin",
                 "particles_00001.bin"):
        assert (out / name).exists(), name
    rows = (out / "summary.csv").read_text().strip().splitlines()
    assert len(rows) == 4  # header + three frames

    # the same loop through the library
    h = kpt.default_host_settings()
    f0 = kpt.compact_field_read(str(fields / "field00000.bin"))
    s = kpt.make_device_settings(h, f0[0], f0[1])
    s.push_chunks = 8
    with kpt.Context(cap, seed=seed) as c:
        for slot, f in ((kpt.SLOT_LOWER, 0), (kpt.SLOT_UPPER, 1)):
            xe, ye, planes = kpt.compact_field_read(str(fields / f"field{f:05d}.bin"))
            c.field_upload(slot, 16, xe, ye, planes)
        for frame in range(3):
            if frame > 0:
                c.field_advance()
                xe, ye, planes = kpt.compact_field_read(str(fields / f"field{frame + 1:05d}.bin"))
                c.field_upload(kpt.SLOT_UPPER, 16, xe, ye, planes)
            c.inject(s, ppf, frame * s.mhd_dt, frame)
            c.push_interval(s, frame * s.mhd_dt, frame)
            c.split(s)
        want = c.download()
        hist = c.histogram_logp(128, h.pmin, h.pmax)
        summ = c.summary()
    snap = kpt.snapshot_read_host(str(out / "particles_00003.bin"), h.charge, h.split_ratio)
    n = len(want["x"])
    for k in ("x", "y", "p", "mu", "weight", "status", "split_level"):
        np.testing.assert_array_equal(snap[k][:n], want[k], err_msg=k)
    ref_csv = tmp_path / "ref_hist.csv"
    kpt.write_histogram_csv(str(ref_csv), h.pmin, h.pmax, hist)
    assert (out / "momentum_histogram_00003.csv").read_bytes() == ref_csv.read_bytes()
    last = rows[-1].split(",")
    assert int(last[0]) == 3 and int(last[2]) == summ["stored"] and int(last[3]) == summ["active"]

    # a fresh run into the same directory is refused; a restart appends
    r = _run(*common, "--frames", "1")
    assert r.returncode == 1 and "already contains reconnection solver outputs" in r.stderr
    r = _run(*common, "--restart-particle-snapshot", str(out / "particles_00003.bin"), "--frames", "1")
    assert r.returncode == 0, r.stderr
    assert "Restarted from particle snapshot" in (out / "run.log").read_text()
    rows2 = (out / "summary.csv").read_text().strip().splitlines()
    assert len(rows2) == 5 and rows2[-1].split(",")[0] == "4"
    assert (out / "particles_00004.bin").exists()


@pytest.mark.gpu
def test_resumable_restart_continues_bit_for_bit(app, tmp_path):
    """Frames 0-4 in one run and frames 0-2 + restart 2-4 from a resumable
    snapshot end in byte-identical particles_00004.bin (ids and split baselines
    travel in the v4 records; the random streams are keyed by seed, id, frame and
    sub-step). A driver-format snapshot of the same state differs only in those
    two fields (checked by the driver-format tests)."""
    fields = tmp_path / "fields"
    fields.mkdir()
    _write_fields(str(fields), 6)
    base = ["--profile", "smoke", "--field-dir", str(fields), "--particles-per-frame", "1500", "--capacity", "20000",
            "--seed", "11", "--particle-snapshot-interval", "2", "--resumable-snapshots", "--push-chunks", "4"]
    a, b = tmp_path / "a", tmp_path / "b"
    r = _run(*base, "--output-dir", str(a), "--frames", "4")
    assert r.returncode == 0, r.stderr
    r = _run(*base, "--output-dir", str(b), "--restart-particle-snapshot", str(a / "particles_00002.bin"),
             "--frames", "2")
    assert r.returncode == 0, r.stderr
    want = (a / "particles_00004.bin").read_bytes()
    got = (b / "particles_00004.bin").read_bytes()
    assert len(want) == len(got) and want == got
    h = kpt.default_host_settings()
    snap = kpt.snapshot_read_host(str(a / "particles_00004.bin"), h.charge, h.split_ratio)
    # the id field carries the particle ids (the cell sort moved them off their slots)
    assert not np.array_equal(snap["id"], np.arange(len(snap["id"]), dtype=np.uint64))


@pytest.mark.gpu
def test_stretched_fields_need_the_flag_and_match_library_loop(app, tmp_path):
    """Compact fields with tanh-stretched y faces: refused with the reference
    reader's message by default; with --allow-stretched-fields the run's particles
    equal a loop driven through the library on the same faces."""
    fields = tmp_path / "fields"
    fields.mkdir()
    ye = W.stretched_edges(32)
    for f in range(3):
        xe, ye, planes = W.island_field(64, 32, width=0.05, seed=3 + f, y_edges=ye)
        planes[2] = planes[2] + 0.1
        kpt.compact_field_write(str(fields / f"field{f:05d}.bin"), xe, ye, planes)
    ppf, cap, seed = 2048, 8192, 5
    common = ["--profile", "smoke", "--field-dir", str(fields), "--particles-per-frame", str(ppf),
              "--capacity", str(cap), "--seed", str(seed), "--frames", "2"]
    r = _run(*common, "--output-dir", str(tmp_path / "refused"))
    assert r.returncode == 1 and "uniformly spaced" in r.stderr, r.stderr
    out = tmp_path / "out"
    r = _run(*common, "--output-dir", str(out), "--allow-stretched-fields")
    assert r.returncode == 0, r.stderr

    h = kpt.default_host_settings()
    xe, ye, _ = kpt.compact_field_read(str(fields / "field00000.bin"), allow_nonuniform=True)
    s = kpt.make_device_settings(h, xe, ye)
    assert s.interpolation_shift_y == 0.0
    s.push_chunks = 8
    with kpt.Context(cap, seed=seed) as c:
        for frame in range(2):
            if frame == 0:
                c.field_upload(kpt.SLOT_LOWER, 16, *kpt.compact_field_read(str(fields / "field00000.bin"),
                                                                           allow_nonuniform=True))
            else:
                c.field_advance()
            c.field_upload(kpt.SLOT_UPPER, 16, *kpt.compact_field_read(str(fields / f"field{frame + 1:05d}.bin"),
                                                                       allow_nonuniform=True))
            c.inject(s, ppf, frame * s.mhd_dt, frame)
            c.push_interval(s, frame * s.mhd_dt, frame)
            c.split(s)
        want = c.download()
    snap = kpt.snapshot_read_host(str(out / "particles_00002.bin"), h.charge, h.split_ratio)
    n = len(want["x"])
    assert n > ppf
    for k in ("x", "y", "p", "weight", "status", "split_level"):
        np.testing.assert_array_equal(snap[k][:n], want[k], err_msg=k)


@pytest.mark.gpu
def test_staged_field_reads_equal_synchronous_reads(app, tmp_path):
    """The default frame loop stages the next field during each push; the run's
    snapshots and histograms equal those of --synchronous-field-reads byte for
    byte."""
    fields = tmp_path / "fields"
    fields.mkdir()
    _write_fields(str(fields), 6)
    base = ["--profile", "smoke", "--field-dir", str(fields), "--particles-per-frame", "1500", "--capacity", "20000",
            "--seed", "3", "--frames", "5", "--particle-snapshot-interval", "2", "--histogram-interval", "2"]
    a, b = tmp_path / "staged", tmp_path / "sync"
    r = _run(*base, "--output-dir", str(a))
    assert r.returncode == 0, r.stderr
    assert "field_reads=staged" in (a / "run.log").read_text()
    r = _run(*base, "--output-dir", str(b), "--synchronous-field-reads")
    assert r.returncode == 0, r.stderr
    names = sorted(p.name for p in a.iterdir() if p.suffix in (".bin", ".csv") and p.name != "summary.csv")
    assert "particles_00005.bin" in names and "momentum_histogram_00004.csv" in names
    for name in names:
        assert (a / name).read_bytes() == (b / name).read_bytes(), name
