#!/usr/bin/env python3
"""Benchmark of the SDE particle push on one MI355X.

One bench step = one 0.1 MHD frame interval of the whole particle set, with the
inputs resident in HBM.

Workload `focused` (default; the headline): focused transport on the
1024 x 512 island-chain Harris field (two time-blended frames; y faces
tanh-stretched toward y = 0 with spacing ratio 10, or uniform with
--grid uniform), 8 388 608
particles with mu ~ U[-1, 1] and p log-uniform on [1, 1e3], slab-QLT
pitch-angle scattering (gamma_k = 5/3, sigma2 = 0.1, Lc = 0.02), fixed
dt = 5e-5 (2000 sub-steps per interval). The measured path is the FP32 bulk
push with the learned D_mumu closure (--closure mlp; the network is trained on
the GPU before the timed region) or with the exact formula (--closure qlt). The
same particles are then pushed by the FP32 exact-QLT push (when the measured
path is the learned one) and by the plain FP64 exact-QLT push, for the ratios.

Workload `parker`: Parker transport at production shape (the same field and
grid), 4 194 304 particles, the driver's default physics with adaptive dt, FP64
or FP32.

Workload `ensemble`: 64 independent runs, each on its own seeded 512 x 256
island-chain field variant (guide field Bz ~ U[0, 0.5], inflow ~ U[0.05, 0.15],
kperp/kpar ~ U[0.01, 0.2]), particle counts log-uniform on [1e3, 1e6]; even runs
Parker, odd runs focused transport with the learned closure; FP32 bulk, fixed
dt = mhd_dt / 50. One bench step is a cell sort of every run followed by a
500-sub-step frame loop (10 intervals) of every run: one persistent launch per
interval, the loop replayed from a hipGraph, the launch shape chosen by the
autotuner first.

Workload `generic`: the generic global-dt solvers (--solver parker: ParkerSolver,
--solver focused: FocusTransportSolver) on the 1024 x 512 island field plus a
0.2 guide field (uniform, periodic), 4 194 304 particles; one bench step is 20
solver steps, each a global dt reduction over all particles and the stochastic
step, FP64, on particles cell-sorted once before the steps. Its roofline prices both passes and takes the time per solver step
from the host clock around the synchronous calls.

Prints one JSON line. `value` is particle-steps/s over the timed steps (exact
sub-step counts from the kernel). `roofline` prices the push kernel's algorithmic
bytes (frames x 4 corners x components x word size per sub-step, plus the
particle state read and written once per launch) against 8 TB/s, with the kernel
time from HIP events recorded on the library's stream around every launch.
`cpu_baseline` times the FP64 CPU restatement of the same push on a bounded
sample of the same particles on this host.
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
import workloads as W  # noqa: E402

METRIC = "particle-steps/s, FP32+MLP-D_μμ vs FP64 plain push, at matched spectrum χ² error"
HBM_PEAK_GBS = 8000.0
GENERIC_STEPS = 20  # solver steps per bench step of the generic workload


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def step_bytes(model, fp32, frames=2, stretched=False):
    """Algorithmic bytes of one sub-step: the coefficient stencil the mover reads,
    plus, on a stretched axis, the bin-table entry and the two faces of the
    particle's interval."""
    return frames * 4 * model * (4 if fp32 else 8) + (4 + 2 * 8 if stretched else 0)


def state_bytes(model):
    """Particle state read and written once per launch."""
    return 88 if model == kpt.MODEL_PARKER16 else 104


def cpu_baseline(xe, ye, f0, f1, model, s, start, target_seconds=10.0):
    """Time the oracle (FP64 restatement of the reference push) on a bounded
    sample of the bench's own particles, using this host's cores."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import kpt_oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = max(1, min(threads, os.cpu_count() or 1))
    lo = kpt_oracle.Field(xe, ye, f0, model)
    hi = kpt_oracle.Field(xe, ye, f1, model)
    s64 = s.copy()
    s64.precision = kpt.PREC_FP64
    if s64.closure == kpt.CLOSURE_MLP:
        s64.closure = kpt.CLOSURE_QLT

    def run(m):
        parts = {k: v[:m].copy() for k, v in start.items()}
        t0 = time.perf_counter()
        steps, _ = kpt_oracle.push_interval(lo, hi, s64, 0.0, 0, 7, parts, threads=threads)
        return steps, time.perf_counter() - t0

    # grow the sample until it takes about target_seconds (small samples scale
    # poorly over the threads, so one extrapolation undershoots)
    m = 1024
    steps, dt = run(m)
    for _ in range(4):
        if dt >= 0.6 * target_seconds or m >= len(start["x"]):
            break
        per_particle_s = dt / m
        m = int(min(len(start["x"]), max(2 * m, target_seconds / max(per_particle_s, 1e-12))))
        steps, dt = run(m)
    what = ("src/main.cpp:1082-1327" if model == kpt.MODEL_PARKER16 else
            "src/main.cpp:1332-1755 with the slab-QLT closure of include/LegencyModel.hpp:539-689 "
            "and the Picard mu update of include/FocusSolver.hpp:700-719")
    return {
        "value": steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{m} of the bench particles, one 0.1 interval ({steps} sub-steps, {dt:.1f} s), "
                  f"FP64 restatement of {what}, OpenMP",
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


def timed_push(ctx, s, warmup, steps, first_frame=0):
    """Warm-up and timed intervals; returns (sub-steps, failures, elapsed s, kernel ms
    per interval, (launches, particles active at entry) per interval)."""
    frame = first_frame
    for _ in range(warmup):
        st = ctx.push_interval(s, frame * s.mhd_dt, frame)
        log(f"  warmup frame {frame}: {st.particle_steps} steps, {st.kernel_ms:.1f} ms")
        frame += 1
    total, fails, kernel_ms, active = 0, 0, [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = ctx.push_interval(s, frame * s.mhd_dt, frame)
        total += st.particle_steps
        fails += st.failures
        kernel_ms.append(st.kernel_ms)
        active.append((st.launches, st.active_in))
        log(f"  frame {frame}: {st.particle_steps} steps, {st.kernel_ms:.1f} ms")
        frame += 1
    elapsed = time.perf_counter() - t0
    return total, fails, elapsed, kernel_ms, active


def profiled_traffic(workload):
    """HBM bytes per push launch of this workload from the committed rocprofv3
    counter summaries (lines `traffic_bytes_per_launch <workload> <bytes>`), or None."""
    import glob

    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                parts = line.strip().strip("`").split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == workload:
                    return float(parts[2])
    return None


def run_generic(args):
    """The generic global-dt solvers (ParkerSolver / FocusTransportSolver): one
    bench step is `GENERIC_STEPS` solver steps of every particle, each a global dt
    reduction over all particles followed by the stochastic step."""
    focus = args.solver == "focused"
    n = args.particles or 4194304
    xe, ye, planes = W.island_field(args.nx, args.ny, seed=7)
    planes[2] = planes[2] + 0.2  # guide field: |B| > 0 everywhere (the solvers abort at nulls)
    if focus:
        p = kpt.generic_focus_default_params()
        p.model.perpendicular_model, p.model.perpendicular_parallel_ratio = 1, 0.05
        p.model.minimum_solar_wind_speed = 1e-3
        p.model.sigma2, p.model.lslab, p.model.l2d, p.model.lc = 0.1, 0.05, 0.02, 0.02
        p.maximum_pitch_angle_cosine = 0.995
        comps = 27
    else:
        p = kpt.generic_default_params()
        p.perpendicular_model, p.perpendicular_parallel_ratio = 1, 0.05
        p.minimum_solar_wind_speed = 1e-3
        p.sigma2, p.lslab, p.l2d = 0.3, 0.05, 0.02
        comps = 18
    rng = np.random.default_rng(3)
    parts = kpt.empty_particles(n)
    parts["x"][:] = rng.uniform(xe[0], xe[-1], n)
    parts["y"][:] = rng.uniform(ye[0], ye[-1], n)
    parts["p"][:] = np.exp(rng.uniform(np.log(0.5), np.log(50.0), n))
    parts["mu"][:] = rng.uniform(-1, 1, n)
    parts["id"][:] = np.arange(n)
    k = GENERIC_STEPS
    log(f"bench: generic {args.solver} solver, {args.nx}x{args.ny}, {n} particles, {k} global-dt steps per step")
    with kpt.Context(n, seed=114514) as ctx:
        ctx.field_upload(kpt.SLOT_LOWER, kpt.MODEL_PARKER16, xe, ye, planes)
        if focus:
            ctx.focus_build(p, kpt.SLOT_LOWER)
        else:
            ctx.generic_build(p, kpt.SLOT_LOWER)
        ctx.upload(parts)
        # ParticleSystem::sort_by_key before the steps, as a production loop does:
        # the stencil gathers of a wavefront then share cache lines (particles move
        # far less than a cell over the timed steps)
        ctx.sort(kpt.make_device_settings(kpt.default_host_settings(), xe, ye))
        step_fn = ctx.focus_step if focus else ctx.generic_step
        step = 0
        fails = 0

        def one_bench_step():
            nonlocal step, fails
            for _ in range(k):
                _, nf = step_fn(p, step)
                fails += nf
                step += 1

        for w in range(args.warmup):
            one_bench_step()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            one_bench_step()
        elapsed = time.perf_counter() - t0
        active = int(np.sum(ctx.download()["status"] == 0))
    # each solver step runs two passes over the particles, every pass reading the
    # 4-corner stencil of the FP64 basis and the particle state
    state_dt = 4 + 24 + (8 if focus else 0)                  # status, x, y, p (, mu)
    state_step = state_dt + 8 + 24 + (8 if focus else 0)      # + id, written x, y, p (, mu)
    per_step = 2 * 4 * comps * 8 + state_dt + state_step
    steps_total = n * k * args.steps
    value = steps_total / elapsed
    secs_per_solver_step = elapsed / (k * args.steps)
    achieved = n * per_step / secs_per_solver_step / 1e9
    workload = f"generic_{args.solver}_{args.nx}x{args.ny}_island_{n}"
    out = {
        "metric": f"particle-steps/s, generic {args.solver} solver (global dt), FP64",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded island-chain Harris field plus a 0.2 guide field; uniform x, y, mu, "
                "log-uniform p)",
        "config": {"workload": workload, "solver": args.solver, "field": f"{args.nx}x{args.ny} uniform, periodic",
                   "particles": n, "solver_steps_per_bench_step": k, "failures": int(fails),
                   "active_after": active, "cell_sorted": True},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(workload),
            "kernel": ("focus" if focus else "generic") + "_dt_kernel + " + ("focus" if focus else "generic")
                      + "_step_kernel",
            "timing": "host clock per synchronous solver step (both launches and the dt read-back)",
            "ms_per_solver_step": secs_per_solver_step * 1e3,
            "algorithmic_bytes_per_solver_step": n * per_step,
            "bytes_per_particle_step": per_step,
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import kpt_oracle

        field = (kpt_oracle.GenericFocusField(xe, ye, planes, p) if focus
                 else kpt_oracle.GenericField(xe, ye, planes, p))
        m, done, t_cpu, s_cpu = 16384, 0, 0.0, 0
        sample = {key: v[:m].copy() for key, v in parts.items()}
        while t_cpu < 10.0 and s_cpu < 2000:
            t1 = time.perf_counter()
            field.step(sample, s_cpu, seed=1)
            t_cpu += time.perf_counter() - t1
            s_cpu += 1
            done += m
        out["cpu_baseline"] = {"value": done / t_cpu, "unit": "particle-steps/s", "cores": 1, "kind": "port",
                               "sample": f"{m} of the bench particles, {s_cpu} global-dt steps ({t_cpu:.1f} s), "
                                         "serial FP64 restatement of the solver's advance_stochastic"}
    print(json.dumps(out), flush=True)


def run_ensemble(args):
    """The batched-ensemble workload (see the module docstring)."""
    specs = W.c5_runs(runs=args.runs or W.C5["runs"])
    frames = W.C5["frames"]
    settings = [W.c5_settings(spec) for spec in specs]
    total = sum(spec["count"] for spec in specs)
    log(f"bench: ensemble of {len(specs)} runs, {total} particles, {frames} x {W.C5['steps_per_interval']} "
        f"sub-steps per step")
    with kpt.Context(16, seed=114514) as ctx:
        log("training the D_mumu network on the GPU ...")
        o = kpt.mlp_default_train_opts()
        o.steps = args.mlp_steps
        o.hull.activation = kpt.MLP_ACTIVATIONS[args.mlp_activation]
        report = ctx.mlp_train(o)
        with kpt.Ensemble(ctx, [spec["count"] for spec in specs]) as e:
            for r, (spec, s) in enumerate(zip(specs, settings)):
                model = kpt.MODEL_FOCUSED21 if spec["focused"] else kpt.MODEL_PARKER16
                e.field_upload(r, model, spec["xe"], spec["ye"], spec["planes"])
                e.set_run(r, s, spec["seed"])
                e.upload(r, W.c5_particles(spec, s, frame=0))
            e.sort()
            chosen, table = e.autotune(0, window=0.2)
            log(f"  autotune: chosen {chosen}; tried {table}")
            every = max(1, min(args.sort_every, frames))

            def loop(frame):
                """A cell sort of every run, then `every` intervals from the graph, until `frames`
                intervals are done. Returns (per-run stats summed, kernel ms)."""
                tot, kms = None, 0.0
                for k in range(0, frames, every):
                    e.sort()
                    st, ms = e.push(frame + k, frames=min(every, frames - k), graph=True)
                    kms += ms
                    tot = st if tot is None else [{f: a[f] + b[f] for f in a} for a, b in zip(tot, st)]
                return tot, kms

            frame = 0
            for _ in range(args.warmup):
                st, ms = loop(frame)
                log(f"  warmup frames {frame}..{frame + frames - 1}: {sum(x['particle_steps'] for x in st)} steps, "
                    f"{ms:.1f} ms")
                frame += frames
            steps, fails, kms = 0, 0, 0.0
            per_kind = {"parker": 0, "focused": 0}
            bytes_alg = 0.0
            t0 = time.perf_counter()
            for _ in range(args.steps):
                st, ms = loop(frame)  # the cell sorts are part of the step
                frame += frames
                kms += ms
                for spec, x in zip(specs, st):
                    steps += x["particle_steps"]
                    fails += x["failures"]
                    model = kpt.MODEL_FOCUSED21 if spec["focused"] else kpt.MODEL_PARKER16
                    per_kind["focused" if spec["focused"] else "parker"] += x["particle_steps"]
                    bytes_alg += x["particle_steps"] * step_bytes(model, True, frames=1) + \
                        x["active_in"] * state_bytes(model)
                log(f"  frames {frame - frames}..{frame - 1}: {sum(x['particle_steps'] for x in st)} steps, "
                    f"{ms:.1f} ms")
            elapsed = time.perf_counter() - t0
    value = steps / elapsed
    launches = args.steps * frames
    avg_launch_s = kms / 1e3 / launches
    achieved = bytes_alg / launches / avg_launch_s / 1e9
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
        "dtype": "f32+f64",
        "data": "synthetic (64 seeded island-chain field variants; C1 Parker and C3 focused particle sets)",
        "config": {
            "workload": f"ensemble_{len(specs)}_runs_512x256_fp32_{total}",
            "runs": len(specs),
            "particles": total,
            "sub_steps_per_step": frames * W.C5["steps_per_interval"],
            "sub_steps": int(steps),
            "parker_sub_steps": int(per_kind["parker"]),
            "focused_mlp_sub_steps": int(per_kind["focused"]),
            "failures": int(fails),
            "tuning": chosen,
            "autotune_table": table,
            "graph": True,
            "sort_every_intervals": every,
            "mlp_training": {k: report[k] for k in ("steps", "train_ms", "val_rms_rel_d")},
            "mlp_activation": args.mlp_activation,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(f"ensemble_{len(specs)}_runs_512x256_fp32_{total}"),
            "kernel": "ensemble_kernel",
            "kernel_ms_avg": avg_launch_s * 1e3,
            "launches": int(launches),
            "algorithmic_bytes_per_launch": bytes_alg / launches,
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        parts = []
        for focused in (False, True):
            spec = max((sp for sp in specs if sp["focused"] == focused), key=lambda sp: sp["count"])
            s = W.c5_settings(spec)
            model = kpt.MODEL_FOCUSED21 if focused else kpt.MODEL_PARKER16
            parts.append(cpu_baseline(spec["xe"], spec["ye"], spec["planes"], spec["planes"], model, s,
                                      W.c5_particles(spec, s, frame=0), target_seconds=6.0))
        share = [per_kind["parker"] / max(steps, 1), per_kind["focused"] / max(steps, 1)]
        rate = 1.0 / sum(w / b["value"] for w, b in zip(share, parts))  # the measured mix of sub-steps
        out["cpu_baseline"] = {"value": rate, "unit": "particle-steps/s", "cores": parts[0]["cores"],
                               "kind": "port",
                               "sample": f"the largest Parker run and the largest focused run (exact QLT), combined "
                                         f"with the measured shares {share[0]:.3f} / {share[1]:.3f} of sub-steps: "
                                         + " | ".join(b["sample"] for b in parts)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run device-to-device copies of 512 MiB (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["focused", "parker", "ensemble", "generic"], default="focused")
    ap.add_argument("--solver", choices=["parker", "focused"], default="parker",
                    help="generic workload: ParkerSolver or FocusTransportSolver")
    ap.add_argument("--runs", type=int, default=0, help="ensemble: runs (0: 64)")
    ap.add_argument("--sort-every", type=int, default=1,
                    help="ensemble: intervals between batched cell sorts of every run")
    ap.add_argument("--particles", type=int, default=0, help="0: the workload's size")
    ap.add_argument("--grid", choices=["stretched", "uniform"], default="stretched",
                    help="focused / parker: y faces of the field")
    ap.add_argument("--stretch-ratio", type=float, default=10.0,
                    help="stretched grid: largest over smallest y spacing")
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default=None,
                    help="default: fp32 for focused, fp64 for parker")
    ap.add_argument("--closure", choices=["qlt", "mlp"], default="mlp")
    ap.add_argument("--mlp-steps", type=int, default=8000, help="Adam steps of the network training")
    ap.add_argument("--mlp-activation", choices=sorted(kpt.MLP_ACTIVATIONS), default="hard-swish",
                    help="hidden-layer activation of the learned closure")
    ap.add_argument("--mlp-passes", type=int, choices=[1, 3], default=1,
                    help="network evaluations per sub-step: 1 (the Picard passes extend the first "
                         "evaluation) or 3 (one per Picard pass)")
    ap.add_argument("--no-fp64-leg", action="store_true",
                    help="focused: skip the plain FP64 push of the same particles")
    ap.add_argument("--no-fp32-leg", action="store_true",
                    help="focused with --closure mlp: skip the FP32 exact-QLT push of the same particles")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--chunks", type=int, default=16,
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

    if args.workload == "ensemble":
        run_ensemble(args)
        return
    if args.workload == "generic":
        run_generic(args)
        return
    focused = args.workload == "focused"
    prec = args.precision or ("fp32" if focused else "fp64")
    fp32 = prec == "fp32"
    model = kpt.MODEL_FOCUSED21 if focused else kpt.MODEL_PARKER16
    stretched = args.grid == "stretched"
    y_edges = W.stretched_edges(args.ny, ratio=args.stretch_ratio) if stretched else None
    xe, ye, f0 = W.island_field(args.nx, args.ny, seed=7, y_edges=y_edges)
    _, _, f1 = W.island_field(args.nx, args.ny, seed=8, y_edges=y_edges)
    if focused:
        n = args.particles or W.C4["particles"]
        closure = kpt.CLOSURE_MLP if args.closure == "mlp" else kpt.CLOSURE_QLT
        s = W.c3_device_settings(xe, ye, closure=closure,
                                 precision=kpt.PREC_FP32 if fp32 else kpt.PREC_FP64)
        s.mlp_passes = args.mlp_passes
        start = W.c3_particles(n, s, seed=11)
    else:
        n = args.particles or 4 * 1024 * 1024
        h = kpt.default_host_settings()
        h.split_particles = 0
        s = kpt.make_device_settings(h, xe, ye)
        s.precision = kpt.PREC_FP32 if fp32 else kpt.PREC_FP64
        start = None
    s.push_chunks = args.chunks
    log(f"bench: {args.workload} {args.nx}x{args.ny} {args.grid}, {n} particles, {prec}"
        + (f", closure {args.closure}" if focused else ""))

    fp64_leg = fp32_leg = mlp_report = None
    with kpt.Context(n, seed=114514) as ctx:
        if focused and args.closure == "mlp":
            log("training the D_mumu network on the GPU ...")
            o = kpt.mlp_default_train_opts()
            o.steps = args.mlp_steps
            o.hull.activation = kpt.MLP_ACTIVATIONS[args.mlp_activation]
            t_train = time.perf_counter()
            mlp_report = ctx.mlp_train(o)
            mlp_report["wall_s"] = time.perf_counter() - t_train
            log(f"  {mlp_report}")
        ctx.field_upload(kpt.SLOT_LOWER, model, xe, ye, f0)
        ctx.field_upload(kpt.SLOT_UPPER, model, xe, ye, f1)
        if focused:
            ctx.upload(start)
        else:
            opts = kpt.InjectOpts()
            opts.fixed_time = 1
            assert ctx.inject(s, n, 0.0, 0, opts) == n
            start = ctx.download()
        steps, failures, elapsed, kernel_ms, active = timed_push(ctx, s, args.warmup, args.steps)
        summary = ctx.summary()
        if focused and args.closure == "mlp" and not args.no_fp32_leg:
            log("FP32 exact-QLT push of the same particles ...")
            s32 = s.copy()
            s32.closure = kpt.CLOSURE_QLT
            ctx.upload(start)
            st32, f32, el32, km32, _ = timed_push(ctx, s32, args.warmup, args.steps)
            fp32_leg = {"value": st32 / el32, "sub_steps": int(st32), "failures": int(f32),
                        "kernel_ms_avg": float(np.mean(km32))}
        if focused and not args.no_fp64_leg:
            log("plain FP64 push of the same particles ...")
            s64 = s.copy()
            s64.precision = kpt.PREC_FP64
            s64.closure = kpt.CLOSURE_QLT
            ctx.upload(start)
            st64, f64, el64, km64, _ = timed_push(ctx, s64, args.warmup, args.steps)
            fp64_leg = {"value": st64 / el64, "sub_steps": int(st64), "failures": int(f64),
                        "kernel_ms_avg": float(np.mean(km64))}

    value = steps / elapsed
    launches = sum(a[0] for a in active)
    avg_launch_s = float(np.sum(kernel_ms)) / 1e3 / launches
    per_step = step_bytes(model, fp32, stretched=stretched)
    launch_bytes = steps / launches * per_step + float(np.mean([a[1] for a in active])) * state_bytes(model)
    achieved = launch_bytes / avg_launch_s / 1e9
    grid_tag = "_stretched" if stretched else ""
    workload = (f"focused_qlt_{args.closure}_{prec}_{args.nx}x{args.ny}{grid_tag}_island_{n}" if focused else
                f"parker_{prec}_{args.nx}x{args.ny}{grid_tag}_island_{n}")
    traffic = profiled_traffic(workload)
    cfg = {
        "workload": workload,
        "field": (f"{args.nx}x{args.ny}, " + (f"y faces tanh-stretched toward y = 0 (spacing ratio {args.stretch_ratio:g})"
                                              if stretched else "uniform grid") + ", 2 time-blended frames"),
        "particles": n,
        "interval": s.mhd_dt,
        "sub_steps": int(steps),
        "failures": int(failures),
        "active_after": int(summary["active"]),
        "push_chunks": args.chunks,
    }
    if focused:
        cfg["closure"] = args.closure
        cfg["dt"] = s.dt_min
        if mlp_report:
            cfg["mlp_activation"] = args.mlp_activation
            cfg["mlp_training"] = {k: mlp_report[k] for k in ("steps", "train_ms", "final_loss",
                                                          "val_rms_lnr", "val_max_lnr",
                                                          "val_rms_rel_d", "val_rms_y1")}
        if fp32_leg:
            cfg["fp32_exact_qlt"] = fp32_leg
            cfg["ratio_vs_fp32_exact_qlt"] = value / fp32_leg["value"]
        if fp64_leg:
            cfg["fp64_plain"] = fp64_leg
            cfg["speedup_vs_fp64_plain"] = value / fp64_leg["value"]
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
        "dtype": "f32+f64" if fp32 else "f64",
        "data": "synthetic (seeded island-chain Harris field; "
                + ("uniform x, y, mu, log-uniform p)" if focused else "uniform injection)"),
        "config": cfg,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "push_kernel_mlp" if (focused and args.closure == "mlp") else "push_kernel",
            "kernel_ms_avg": avg_launch_s * 1e3,
            "launches": int(launches),
            "algorithmic_bytes_per_launch": launch_bytes,
            "bytes_per_step": per_step,
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(xe, ye, f0, f1, model, s, start)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
