#!/usr/bin/env python3
"""Benchmark of the Parker SDE push on one MI355X.

Workload (one bench step = one MHD frame interval of the whole particle set):
Parker transport at production shape, FP64 plain push: 1024 x 512 Harris sheet
with a 4-mode island chain and 1% velocity noise, two time-blended frames, the
driver's default kappa / drift physics (p0 = 0.1, adaptive dt), 4 194 304
particles resident in HBM, one 0.1 interval per step.

Prints one JSON line. `value` is particle-steps/s over the timed steps (exact
sub-step counts from the kernel). `roofline` prices the push kernel's algorithmic
bytes (2 frames x 4 corners x 16 components x 8 B per sub-step, plus the
particle state read and written once per launch) against 8 TB/s, with the
kernel time measured by HIP events on the library's stream. `cpu_baseline` times
the FP64 CPU restatement of the reference push on a bounded sample of the same
particles on this host.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "kokkos-particletransport_amd"))

import numpy as np  # noqa: E402

import kpt  # noqa: E402
import workloads  # noqa: E402

METRIC = "particle-steps/s, FP32+MLP-D_μμ vs FP64 plain push, at matched spectrum χ² error"
HBM_PEAK_GBS = 8000.0
BYTES_PER_STEP = 2 * 4 * 16 * 8  # frames x corners x components x sizeof(double)
STATE_BYTES = 88                  # particle state read + written per launch


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(nx, ny, n, seed):
    xe, ye, f0 = workloads.island_field(nx, ny, seed=seed)
    _, _, f1 = workloads.island_field(nx, ny, seed=seed + 1)
    h = kpt.default_host_settings()
    h.split_ratio = 1.2
    h.split_particles = 0
    s = kpt.make_device_settings(h, xe, ye)
    return xe, ye, f0, f1, s


def cpu_baseline(xe, ye, f0, f1, s, start, target_seconds=10.0):
    """Time the oracle (FP64 restatement of the reference push) on a bounded
    sample of the bench's own particles, using this host's cores."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import kpt_oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))
    lo = kpt_oracle.Field(xe, ye, f0, kpt.MODEL_PARKER16)
    hi = kpt_oracle.Field(xe, ye, f1, kpt.MODEL_PARKER16)

    def run(m):
        parts = {k: v[:m].copy() for k, v in start.items()}
        t0 = time.perf_counter()
        steps, _ = kpt_oracle.push_interval(lo, hi, s, 0.0, 0, 7, parts, threads=threads)
        return steps, time.perf_counter() - t0

    m = 2048
    steps, dt = run(m)
    rate = steps / max(dt, 1e-9)
    per_particle = steps / m
    m = int(min(len(start["x"]), max(m, target_seconds * rate / max(per_particle, 1.0))))
    steps, dt = run(m)
    return {
        "value": steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{m} of the bench particles, one 0.1 interval ({steps} sub-steps, {dt:.1f} s), "
                  f"FP64 restatement of src/main.cpp:1082-1327 with OpenMP",
    }


def copy_probe(mib=512, repeats=3):
    """Device-to-device copies of a known size, for scaling the FETCH_SIZE /
    WRITE_SIZE counters of a rocprofv3 --pmc run (the copy reads and writes
    exactly `mib` MiB per call)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(repeats):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_mib": mib, "repeats": repeats}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run device-to-device copies of 512 MiB (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=4 * 1024 * 1024)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sort", action="store_true",
                    help="push in injection order instead of sorting by cell before each interval")
    ap.add_argument("--chunks", type=int, default=8,
                    help="time chunks per interval, each preceded by a cell sort (push_chunks)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(particles would shard by index range with one histogram all-reduce per "
              "diagnostic interval; not built).", file=sys.stderr)
        sys.exit(2)

    xe, ye, f0, f1, s = build_workload(args.nx, args.ny, args.particles, seed=7)
    s.precision = kpt.PREC_FP32 if args.precision == "fp32" else kpt.PREC_FP64
    n = args.particles
    log(f"bench: {args.nx}x{args.ny} field, {n} particles, {args.precision}")
    with kpt.Context(n, seed=114514) as ctx:
        ctx.field_upload(kpt.SLOT_LOWER, kpt.MODEL_PARKER16, xe, ye, f0)
        ctx.field_upload(kpt.SLOT_UPPER, kpt.MODEL_PARKER16, xe, ye, f1)
        opts = kpt.InjectOpts()
        opts.fixed_time = 1
        assert ctx.inject(s, n, 0.0, 0, opts) == n
        start = ctx.download()

        s.push_chunks = 1 if args.no_sort else args.chunks

        def interval(frame):
            # one bench step = one interval: with push_chunks > 1 the library
            # re-sorts by cell before every chunk; with one chunk sort here
            if not args.no_sort and s.push_chunks <= 1:
                ctx.sort(s)
            return ctx.push_interval(s, frame * s.mhd_dt, frame)

        frame = 0
        for _ in range(args.warmup):
            st = interval(frame)
            log(f"warmup frame {frame}: {st.particle_steps} steps, {st.kernel_ms:.1f} ms")
            frame += 1
        steps = 0
        kernel_ms = []
        launches_bytes = []
        failures = 0
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = interval(frame)
            steps += st.particle_steps
            failures += st.failures
            kernel_ms.append(st.kernel_ms)
            launches_bytes.append(st.particle_steps * BYTES_PER_STEP + st.active_in * STATE_BYTES)
            log(f"frame {frame}: {st.particle_steps} steps, {st.kernel_ms:.1f} ms")
            frame += 1
        elapsed = time.perf_counter() - t0
        summary = ctx.summary()

    value = steps / elapsed
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    achieved = float(np.mean(launches_bytes)) / avg_kernel_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded island-chain Harris field, uniform injection)",
        "config": {
            "workload": f"parker_{args.precision}_plain_push_{args.nx}x{args.ny}_island_{n}",
            "field": f"{args.nx}x{args.ny} uniform grid, 2 time-blended frames",
            "particles": n,
            "interval": s.mhd_dt,
            "sub_steps": int(steps),
            "failures": int(failures),
            "active_after": int(summary["active"]),
            "path": "FP64 plain Parker push (the FP32+MLP focused path is not built yet)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel_ms_avg": avg_kernel_s * 1e3,
            "bytes_per_step": BYTES_PER_STEP,
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(xe, ye, f0, f1, s, start)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
