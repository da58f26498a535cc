"""The C-ABI library loads without a GPU and exports every function that
include/kpt.h declares; the Python binding declares the same set."""
import os
import re
import subprocess

import kpt

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(REPO, "include", "kpt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:kpt_status|const char\*|void)\s+(kpt_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    fns = header_functions()
    for need in ("kpt_create", "kpt_push_interval", "kpt_inject", "kpt_split", "kpt_histogram_logp",
                 "kpt_summary_compute", "kpt_field_load_compact", "kpt_snapshot_write",
                 "kpt_snapshot_read", "kpt_set_noise"):
        assert need in fns
    assert len(fns) >= 35


def test_library_exports_every_declared_symbol(kpt_mod):
    out = subprocess.run(["nm", "-D", "--defined-only", kpt.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (kpt_\w+)$", out, re.M))
    missing = [f for f in header_functions() if f not in exported]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(kpt.EXPORTS) == header_functions()


def test_library_is_gfx950(kpt_mod):
    """The device code object inside libkpt.so targets gfx950."""
    blob = open(kpt.LIB_PATH, "rb").read()
    assert b"gfx950" in blob
    assert kpt.load().kpt_version().decode().endswith("(gfx950)")


def test_create_without_gpu_fails_cleanly(kpt_mod):
    """Here (no GPU) kpt_create must fail with a message, not crash or fall back."""
    import pytest

    try:
        ctx = kpt.Context(16)
    except kpt.KptError as e:
        assert "gfx950" in str(e) or "device" in str(e)
        return
    ctx.close()
    pytest.skip("a GPU is present; covered by the gpu tests")


def test_noalloc_limits_match_the_header():
    """The no-allocation push's published limits: the binding's constants are the
    header's, and the scratch array fits the chunk records of the launch
    arguments the kernels read."""
    text = open(os.path.join(REPO, "include", "kpt.h")).read()
    consts = dict(re.findall(r"^#define (KPT_NOALLOC_\w+) (\d+)$", text, re.M))
    assert int(consts["KPT_NOALLOC_MAX_CHUNKS"]) == kpt.NOALLOC_MAX_CHUNKS
    assert int(consts["KPT_NOALLOC_SCRATCH_BYTES"]) == kpt.NOALLOC_SCRATCH_BYTES
    src = open(os.path.join(REPO, "kokkos-particletransport_amd", "csrc", "kpt_noalloc.hip")).read()
    assert re.search(r"^static __device__ __align__\(256\) unsigned char scratch\[KPT_NOALLOC_SCRATCH_BYTES\];",
                     src, re.M)
    assert "hipMalloc" not in re.sub(r"//.*", "", src) and "hipFree" not in re.sub(r"//.*", "", src)
