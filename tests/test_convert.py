"""kpt_convert_mhd, the Fortran MHD -> compact-field converter
(apps/fortran_mhd_to_compact_field.cpp:171-250): on synthetic Fortran frames
(config of 13 f64 + 13 i32; f32 [ny + 4][nx + 4][8] with two ghost layers) the
output equals, byte for byte, an independent numpy conversion (analytic faces,
planes Bx, By, Bz, Vx, Vy, Vz of the physical cells), the library's reader
accepts it, and bad inputs end with the converter's messages. No GPU."""
import os
import subprocess

import numpy as np
import pytest

import kpt

HERE = os.path.dirname(os.path.abspath(__file__))
TOOL = os.path.join(os.path.dirname(HERE), "kokkos-particletransport_amd", "kpt_convert_mhd")


@pytest.fixture(scope="module")
def tool():
    if not os.path.exists(TOOL):
        pytest.fail(f"{TOOL} is missing: build it with `make -C kokkos-particletransport_amd`")
    return TOOL


def _config(path, nx, ny, nz=1, lo=(-1.0, -0.5), hi=(1.0, 0.5)):
    real = np.array([0.1, 0.1, 1.0, lo[0], lo[1], 0.0, hi[0], hi[1], 1.0, 2.0, 1.0, 1.0, 0.1], dtype="<f8")
    ints = np.array([nx, ny, nz, nx, ny, 1, 1, 1, 1, 8, 0, 0, 0], dtype="<i4")
    with open(path, "wb") as f:
        f.write(real.tobytes())
        f.write(ints.tobytes())


def _expected(raw, nx, ny, lo, hi):
    out = np.array([nx, ny], dtype="<i4").tobytes()
    for d, n in enumerate((nx, ny)):
        step = (hi[d] - lo[d]) / n
        out += np.array([lo[d] + k * step for k in range(n + 1)], dtype="<f8").tobytes()
    for v in (4, 5, 6, 0, 1, 2):
        out += raw[2:2 + ny, 2:2 + nx, v].astype("<f8").tobytes()
    return out


def test_conversion_matches_an_independent_restatement(tool, tmp_path):
    nx, ny, lo, hi = 24, 10, (-1.0, -0.5), (1.0, 0.5)
    src, dst = tmp_path / "mhd", tmp_path / "compact"
    src.mkdir()
    _config(src / "mhd_config.dat", nx, ny, lo=lo, hi=hi)
    rng = np.random.default_rng(4)
    raws = []
    for f in range(3, 6):
        raw = rng.normal(0.0, 1.0, (ny + 4, nx + 4, 8)).astype("<f4")
        (src / f"mhd_data_{f:04d}").write_bytes(raw.tobytes())
        raws.append(raw)
    r = subprocess.run([tool, "--input-dir", str(src), "--output-dir", str(dst), "--start-frame", "3",
                        "--end-frame", "5"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "frame 5 ->" in r.stdout
    for f, raw in zip(range(3, 6), raws):
        got = (dst / f"field{f:05d}.bin").read_bytes()
        assert got == _expected(raw, nx, ny, lo, hi)
        xe, ye, planes = kpt.compact_field_read(str(dst / f"field{f:05d}.bin"))
        np.testing.assert_array_equal(planes[0], raw[2:2 + ny, 2:2 + nx, 4].astype(np.float64))  # Bx
        np.testing.assert_array_equal(planes[4], raw[2:2 + ny, 2:2 + nx, 1].astype(np.float64))  # Vy


def test_converter_errors(tool, tmp_path):
    src = tmp_path / "mhd"
    src.mkdir()
    run = lambda *a: subprocess.run([tool, *a], capture_output=True, text=True, timeout=60)  # noqa: E731
    r = run("--input-dir", str(src), "--output-dir", str(tmp_path / "o"))
    assert r.returncode == 1 and "failed to open MHD config" in r.stderr
    _config(src / "mhd_config.dat", 8, 8, nz=4)
    r = run("--input-dir", str(src), "--output-dir", str(tmp_path / "o"))
    assert r.returncode == 1 and "only positive 2D MHD configurations are supported" in r.stderr
    _config(src / "mhd_config.dat", 8, 8)
    (src / "mhd_data_0000").write_bytes(b"\0" * 100)
    r = run("--input-dir", str(src), "--output-dir", str(tmp_path / "o"), "--end-frame", "0")
    assert r.returncode == 1 and "failed to read MHD frame" in r.stderr
    r = run("--input-dir", str(src), "--output-dir", str(tmp_path / "o"), "--start-frame", "3", "--end-frame", "2")
    assert r.returncode == 1 and "end frame must be greater than or equal to start frame" in r.stderr
    r = run("--bogus")
    assert r.returncode == 1 and "unknown option: --bogus" in r.stderr
