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

Workload `parker`: the production Parker loop on the same field and grid
as a drifting frame series, the driver's default physics with adaptive dt, FP64
or FP32. One bench step is one frame of the reference's loop: the upper frame is
uploaded, 1600 x 16 = 25 600 particles are injected, every stored particle is
pushed through the interval (16 cell-sorted chunks), particles split at
thresholds x1.2 per level from 2 p0, and the momentum histogram is taken every
10 frames (the reference's production script, scripts/run_total_benchmark_suite_
local.sh:132-134). The timed window starts at frame 160 (--loop-start-frame) with
the 160 x 25 600 particles of the earlier frames stored, in their injected state
(uniform in the domain at the window's start time, p = p0), so a short run
measures the loop at a late-run particle count. --parker-push-only pushes 4 194 304 particles injected once instead.

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


def host_cpu():
    """This host's CPU model and logical CPU count (the CPU baseline's machine)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


def cpu_baseline(xe, ye, f0, f1, model, s, start, target_seconds=10.0, seed=7):
    """Time the oracle (FP64 restatement of the reference push) on a bounded
    sample of the bench's own particles, using this host's cores. Returns the
    baseline record, the sample's start state and the oracle's end state."""
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
        steps, _ = kpt_oracle.push_interval(lo, hi, s64, 0.0, 0, seed, parts, threads=threads)
        return steps, time.perf_counter() - t0, parts

    # grow the sample until it takes about target_seconds (small samples scale
    # poorly over the threads, so one extrapolation undershoots)
    m = 1024
    steps, dt, end = run(m)
    for _ in range(4):
        if dt >= 0.6 * target_seconds or m >= len(start["x"]):
            break
        per_particle_s = dt / m
        m = int(min(len(start["x"]), max(2 * m, target_seconds / max(per_particle_s, 1e-12))))
        steps, dt, end = run(m)
    what = ("src/main.cpp:1082-1327" if model == kpt.MODEL_PARKER16 else
            "src/main.cpp:1332-1755 with the slab-QLT closure of include/LegencyModel.hpp:539-689 "
            "and the Picard mu update of include/FocusSolver.hpp:700-719")
    rec = {
        "value": steps / dt,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "host": host_cpu(),
        "sample": f"{m} of the bench particles, one 0.1 interval ({steps} sub-steps, {dt:.1f} s), "
                  f"FP64 restatement of {what}, OpenMP",
    }
    return rec, {k: v[:m].copy() for k, v in start.items()}, end


# MI355X peaks of the compute-side roofline: 256 CUs x 4 SIMDs at 2.4 GHz; a
# wave64 VALU instruction holds its SIMD for 4 cycles; unpacked FP32 and FP64 FMA
# 78.6 TFLOP/s (packed FP32 is off in this build); dense BF16 MFMA 2.5 PFLOP/s.
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4
TA_BUSY_PEAK = 256 * 2.4e9  # one texture address unit per CU at the 2.4 GHz engine clock
FP32_PEAK = 78.6e12
FP32_PUBLISHED_PEAK = 157.3e12  # with packed FP32 (two lanes per instruction)
FP64_PEAK = 78.6e12
BF16_MFMA_PEAK = 2.5e15


def profiled_counters(workload):
    """Per-launch counters of this workload's dominant kernel from the committed
    rocprofv3 summaries (lines `counters_per_launch <workload> name=value ...`),
    with the summary's file name, or ({}, None)."""
    import glob

    for path in sorted(glob.glob(os.path.join(REPO, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                parts = line.strip().strip("`").split()
                if len(parts) > 2 and parts[0] == "counters_per_launch" and parts[1] == workload:
                    return {k: float(v) for k, v in (p.split("=", 1) for p in parts[2:])}, os.path.basename(path)
    return {}, None


def roofline_block(workload, kernel, launch_s, launches, alg_bytes):
    """The roofline of the dominant kernel on every resource the profile
    measures, with `bound` the one nearest its peak. Counters per launch come
    from profiles/ (instruction and FLOP counts, HBM bytes); times are this run's
    HIP-event launch averages."""
    c, src = profiled_counters(workload)
    traffic = profiled_traffic(workload)
    hbm_alg = alg_bytes / launch_s / 1e9
    res = {"hbm_algorithmic": {"achieved": hbm_alg, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": hbm_alg / HBM_PEAK_GBS,
                               "note": "algorithmic stencil bytes; cell-sorted gathers are served by L2 and MALL, "
                                       "so this can exceed 1 and is not HBM evidence"}}
    if traffic is not None:
        g = traffic / launch_s / 1e9
        res["hbm"] = {"achieved": g, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": g / HBM_PEAK_GBS}
    if "SQ_INSTS_VALU" in c:
        v = c["SQ_INSTS_VALU"] / launch_s
        res["valu_issue"] = {"achieved": v / 1e9, "peak": VALU_ISSUE_PEAK / 1e9, "unit": "G wave-instr/s",
                             "frac": v / VALU_ISSUE_PEAK}
    if "SQ_ACTIVE_INST_VALU" in c:
        # VALU pipe occupancy: issue slots of 4 cycles per wave64 instruction, 8 for
        # a transcendental
        v = c["SQ_ACTIVE_INST_VALU"] / launch_s
        res["valu_busy"] = {"achieved": v / 1e9, "peak": VALU_ISSUE_PEAK / 1e9, "unit": "G wave-slots/s",
                            "frac": v / VALU_ISSUE_PEAK,
                            "note": "peak at the nominal 2.4 GHz engine clock: a fraction at or slightly above 1 "
                                    "is a saturated VALU pipe"}
    if "TA_TA_BUSY_sum" in c:
        # vector memory address unit (one per CU) busy cycles: the gather path
        v = c["TA_TA_BUSY_sum"] / launch_s
        res["vmem_ta"] = {"achieved": v / 1e9, "peak": TA_BUSY_PEAK / 1e9, "unit": "G busy cycles/s",
                          "frac": v / TA_BUSY_PEAK}
    if "FLOP_F32" in c:
        v = c["FLOP_F32"] / launch_s
        res["fp32"] = {"achieved": v / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s", "frac": v / FP32_PEAK,
                       "published_peak": FP32_PUBLISHED_PEAK / 1e12,
                       "note": "the published peak counts packed FP32, which this build does not issue"}
    if "FLOP_F64" in c:
        v = c["FLOP_F64"] / launch_s
        res["fp64"] = {"achieved": v / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s", "frac": v / FP64_PEAK}
    if c.get("FLOP_BF16_MFMA", 0.0) > 0.0:
        v = c["FLOP_BF16_MFMA"] / launch_s
        res["bf16_mfma"] = {"achieved": v / 1e12, "peak": BF16_MFMA_PEAK / 1e12, "unit": "TFLOP/s",
                            "frac": v / BF16_MFMA_PEAK}
    measured = {k: r for k, r in res.items() if k != "hbm_algorithmic"}
    bound = max(measured, key=lambda k: measured[k]["frac"]) if measured else "hbm_algorithmic"
    top = res[bound]
    out = {"bound": bound, "achieved": top["achieved"], "peak": top["peak"], "unit": top["unit"],
           "frac": top["frac"], "traffic": traffic, "kernel": kernel, "kernel_ms_avg": launch_s * 1e3,
           "launches": int(launches), "algorithmic_bytes_per_launch": alg_bytes, "resources": res}
    if src:
        out["counters"] = {"source": f"profiles/{src}", "per_launch": c,
                           "profiled_kernel_ms": c.get("kernel_ms"),
                           "stale": bool(c.get("kernel_ms") and abs(c["kernel_ms"] / (launch_s * 1e3) - 1) > 0.1)}
    return out


def accuracy_gate(xe, ye, f0, f1, model, s, net, sample, want, oracle_seed, other_seed=2718):
    """The measured path on the CPU baseline's sample against the oracle's end
    state: with the oracle's seed, the median per-particle differences; with an
    independent seed, two-sample chi-square and KS p-values of the final
    distributions and of the displacements (the metric's matched-spectrum
    condition, passed when every p-value exceeds 0.01 divided by the number of
    tests, a family-wise level of 0.01)."""
    import accuracy as A

    def push(seed):
        with kpt.Context(len(sample["x"]), seed=seed) as c:
            if net is not None:
                c.mlp_set(*net)
            c.field_upload(kpt.SLOT_LOWER, model, xe, ye, f0)
            c.field_upload(kpt.SLOT_UPPER, model, xe, ye, f1)
            c.upload(sample)
            st = c.push_interval(s, 0.0, 0)
            return A.sort_by_id(c.download()), st

    want = A.sort_by_id(want)
    start = A.sort_by_id(sample)
    same, _ = push(oracle_seed)
    other, st = push(other_seed)
    if model == kpt.MODEL_FOCUSED21:
        p = A.focused_pvalues(want, other, start, s)
    else:
        p = A.displacement_pvalues(want, other, start, s.x_width, s.y_width, ("dx", "dy", "dlnp"))
    return {"particles": len(sample["x"]), "failures": int(st.failures),
            "pathwise_median_same_seed": A.pathwise_medians(same, want, s.x_width, s.y_width),
            "median_abs_dmu_oracle": float(np.median(np.abs(want["mu"] - start["mu"]))) if "mu" in want else None,
            "pvalues_independent_seed": p, "min_pvalue": min(p.values()),
            # family-wise level 0.01 over all the tests (Bonferroni)
            "threshold": 0.01 / len(p),
            "passed": bool(min(p.values()) > 0.01 / len(p) and st.failures == 0)}


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
        "roofline": roofline_block(workload, ("focus" if focus else "generic") + "_dt_kernel + "
                                   + ("focus" if focus else "generic") + "_step_kernel", secs_per_solver_step,
                                   k * args.steps, n * per_step),
    }
    out["roofline"]["timing"] = "host clock per synchronous solver step (both launches and the dt read-back)"
    out["roofline"]["bytes_per_particle_step"] = per_step
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import kpt_oracle

        field = (kpt_oracle.GenericFocusField(xe, ye, planes, p) if focus
                 else kpt_oracle.GenericField(xe, ye, planes, p))
        threads = kpt_oracle.max_threads()
        m, done, t_cpu, s_cpu = 131072, 0, 0.0, 0
        sample = {key: v[:m].copy() for key, v in parts.items()}
        while t_cpu < 10.0 and s_cpu < 2000:
            t1 = time.perf_counter()
            field.step(sample, s_cpu, seed=1)
            t_cpu += time.perf_counter() - t1
            s_cpu += 1
            done += m
        out["cpu_baseline"] = {"value": done / t_cpu, "unit": "particle-steps/s", "cores": threads, "kind": "port",
                               "host": host_cpu(),
                               "sample": f"{m} of the bench particles, {s_cpu} global-dt steps ({t_cpu:.1f} s), "
                                         "FP64 restatement of the solver's compute_adaptive_time_step and "
                                         "advance_stochastic, OpenMP over particles"}
    print(json.dumps(out), flush=True)


def run_ensemble(args):
    """The batched-ensemble workload (see the module docstring)."""
    specs = W.ensemble_runs(runs=args.runs or W.ENSEMBLE_64X512X256["runs"])
    frames = W.ENSEMBLE_64X512X256["frames"]
    settings = [W.ensemble_run_settings(spec) for spec in specs]
    total = sum(spec["count"] for spec in specs)
    log(f"bench: ensemble of {len(specs)} runs, {total} particles, {frames} x {W.ENSEMBLE_64X512X256['steps_per_interval']} "
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
                e.upload(r, W.ensemble_run_particles(spec, s, frame=0))
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
    out = {
        "metric": "particle-steps/s, batched ensemble of independent runs (FP32, Parker and learned-closure focused)",
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
        "data": "synthetic (64 seeded island-chain field variants; Harris-sheet Parker and slab-QLT focused particle sets)",
        "config": {
            "workload": f"ensemble_{len(specs)}_runs_512x256_fp32_{total}",
            "runs": len(specs),
            "particles": total,
            "sub_steps_per_step": frames * W.ENSEMBLE_64X512X256["steps_per_interval"],
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
        "roofline": roofline_block(f"ensemble_{len(specs)}_runs_512x256_fp32_{total}", "ensemble_kernel",
                                   avg_launch_s, launches, bytes_alg / launches),
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        parts = []
        for focused in (False, True):
            spec = max((sp for sp in specs if sp["focused"] == focused), key=lambda sp: sp["count"])
            s = W.ensemble_run_settings(spec)
            model = kpt.MODEL_FOCUSED21 if focused else kpt.MODEL_PARKER16
            parts.append(cpu_baseline(spec["xe"], spec["ye"], spec["planes"], spec["planes"], model, s,
                                      W.ensemble_run_particles(spec, s, frame=0), target_seconds=6.0)[0])
        share = [per_kind["parker"] / max(steps, 1), per_kind["focused"] / max(steps, 1)]
        rate = 1.0 / sum(w / b["value"] for w, b in zip(share, parts))  # the measured mix of sub-steps
        out["cpu_baseline"] = {"value": rate, "unit": "particle-steps/s", "cores": parts[0]["cores"],
                               "host": host_cpu(),
                               "kind": "port",
                               "sample": f"the largest Parker run and the largest focused run (exact QLT), combined "
                                         f"with the measured shares {share[0]:.3f} / {share[1]:.3f} of sub-steps: "
                                         + " | ".join(b["sample"] for b in parts)}
    print(json.dumps(out), flush=True)


def _leg(ctx, s, start, warmup, steps):
    ctx.upload(start)
    st, f, el, km, act = timed_push(ctx, s, warmup, steps)
    launches = sum(a[0] for a in act)
    return {"value": st / el, "sub_steps": int(st), "failures": int(f),
            "kernel_ms_avg": float(np.sum(km)) / launches, "launches": int(launches),
            "active_in": float(np.mean([a[1] for a in act]))}


def run_push(args):
    """The focused (headline) and Parker push workloads (see the module docstring)."""
    focused = args.workload == "focused"
    prec = args.precision or ("fp32" if focused else "fp64")
    fp32 = prec == "fp32"
    model = kpt.MODEL_FOCUSED21 if focused else kpt.MODEL_PARKER16
    stretched = args.grid == "stretched"
    y_edges = W.stretched_edges(args.ny, ratio=args.stretch_ratio) if stretched else None
    xe, ye, f0 = W.island_field(args.nx, args.ny, seed=7, y_edges=y_edges)
    _, _, f1 = W.island_field(args.nx, args.ny, seed=8, y_edges=y_edges)
    learned = focused and args.closure == "mlp"
    if focused:
        n = args.particles or W.FOCUSED_QLT_1024X512_8M["particles"]
        closure = kpt.CLOSURE_MLP if learned else kpt.CLOSURE_QLT
        s = W.focused_qlt_device_settings(xe, ye, closure=closure,
                                 precision=kpt.PREC_FP32 if fp32 else kpt.PREC_FP64)
        s.mlp_passes = args.mlp_passes
        start = W.focused_qlt_particles(n, s, seed=11)
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

    legs, reports, nets = {}, {}, {}
    with kpt.Context(n, seed=114514) as ctx:
        ctx.field_upload(kpt.SLOT_LOWER, model, xe, ye, f0)
        ctx.field_upload(kpt.SLOT_UPPER, model, xe, ye, f1)
        if not focused:
            opts = kpt.InjectOpts()
            opts.fixed_time = 1
            assert ctx.inject(s, n, 0.0, 0, opts) == n
            start = ctx.download()
        acts = [args.mlp_activation] if learned else []
        if learned and not args.no_alt_activation_leg:
            acts += [a for a in sorted(kpt.MLP_ACTIVATIONS) if a != args.mlp_activation]
        for act in acts:
            log(f"training the D_mumu network ({act}) on the GPU ...")
            o = kpt.mlp_default_train_opts()
            o.steps = args.mlp_steps
            o.hull.activation = kpt.MLP_ACTIVATIONS[act]
            if args.mlp_mu_lo > 0:
                o.hull.mu_lo = args.mlp_mu_lo
            t_train = time.perf_counter()
            reports[act] = ctx.mlp_train(o)
            reports[act]["wall_s"] = time.perf_counter() - t_train
            nets[act] = ctx.mlp_get()
            log(f"  {reports[act]}")
        # the measured path first, then the comparison legs on the same particles
        if learned:
            for act in acts:
                log(f"FP32 push with the learned closure ({act}) ...")
                ctx.mlp_set(*nets[act])
                legs["mlp_" + act] = _leg(ctx, s, start, args.warmup, args.steps)
            ctx.mlp_set(*nets[args.mlp_activation])
        else:
            legs["measured"] = _leg(ctx, s, start, args.warmup, args.steps)
        main_leg = legs["mlp_" + args.mlp_activation] if learned else legs["measured"]
        summary = ctx.summary()
        if learned and not args.no_fp32_leg:
            log("FP32 exact-QLT push of the same particles ...")
            s32 = s.copy()
            s32.closure = kpt.CLOSURE_QLT
            legs["fp32_exact_qlt"] = _leg(ctx, s32, start, args.warmup, args.steps)
        if focused and not args.no_fp64_leg:
            log("plain FP64 push of the same particles ...")
            s64 = s.copy()
            s64.precision = kpt.PREC_FP64
            s64.closure = kpt.CLOSURE_QLT
            legs["fp64_plain"] = _leg(ctx, s64, start, args.warmup, args.steps)

    steps = main_leg["sub_steps"]
    value = main_leg["value"]
    launch_s = main_leg["kernel_ms_avg"] / 1e3
    per_step = step_bytes(model, fp32, stretched=stretched)
    launch_bytes = steps / main_leg["launches"] * per_step + main_leg["active_in"] * state_bytes(model)
    grid_tag = "_stretched" if stretched else ""

    def name(closure, act, p):
        act_tag = ("_" + act.replace("-", "")) if closure == "mlp" else ""
        return (f"focused_qlt_{closure}{act_tag}_{p}_{args.nx}x{args.ny}{grid_tag}_island_{n}" if focused
                else f"parker_{p}_{args.nx}x{args.ny}{grid_tag}_island_{n}")

    workload = name(args.closure, args.mlp_activation, prec)
    kernel = "push_kernel_mlp" if learned else "push_kernel"
    # each comparison leg's roofline under its own workload name
    for key, leg in legs.items():
        if leg is main_leg:
            continue
        if key.startswith("mlp_"):
            lw, lk, lp = name("mlp", key[4:], "fp32"), "push_kernel_mlp", True
        elif key == "fp32_exact_qlt":
            lw, lk, lp = name("qlt", "", "fp32"), "push_kernel", True
        else:
            lw, lk, lp = name("qlt", "", "fp64"), "push_kernel", False
        lb = leg["sub_steps"] / leg["launches"] * step_bytes(model, lp, stretched=stretched) + \
            leg["active_in"] * state_bytes(model)
        leg["workload"] = lw
        leg["roofline"] = roofline_block(lw, lk, leg["kernel_ms_avg"] / 1e3, leg["launches"], lb)
    cfg = {
        "workload": workload,
        "field": (f"{args.nx}x{args.ny}, " + (f"y faces tanh-stretched toward y = 0 (spacing ratio {args.stretch_ratio:g})"
                                              if stretched else "uniform grid") + ", 2 time-blended frames"),
        "particles": n,
        "interval": s.mhd_dt,
        "sub_steps": int(steps),
        "failures": int(main_leg["failures"]),
        "active_after": int(summary["active"]),
        "push_chunks": args.chunks,
    }
    if focused:
        cfg["closure"] = args.closure
        cfg["dt"] = s.dt_min
    if learned:
        cfg["mlp_activation"] = args.mlp_activation
        cfg["mlp_network"] = "3 -> 64 -> 64 -> 64 -> 2, BF16 weights, FP32 accumulation (MFMA 32x32x16)"
        cfg["mlp_evaluations_per_sub_step"] = 1 if args.mlp_passes == 1 else 3
        cfg["mlp_training"] = {a: {k: reports[a][k] for k in ("steps", "train_ms", "final_loss", "val_rms_lnr",
                                                               "val_max_lnr", "val_rms_rel_d", "val_rms_y1")}
                               for a in reports}
    cfg["legs"] = legs
    if "fp64_plain" in legs:
        cfg["speedup_vs_fp64_plain"] = value / legs["fp64_plain"]["value"]
        for key, leg in legs.items():
            if key.startswith("mlp_"):
                cfg[f"speedup_vs_fp64_plain_{key[4:]}"] = leg["value"] / legs["fp64_plain"]["value"]
        if "fp32_exact_qlt" in legs:
            cfg["speedup_vs_fp64_plain_fp32_exact_qlt"] = legs["fp32_exact_qlt"]["value"] / legs["fp64_plain"]["value"]
    if "fp32_exact_qlt" in legs:
        cfg["ratio_vs_fp32_exact_qlt"] = value / legs["fp32_exact_qlt"]["value"]
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_leg["sub_steps"] / value / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if fp32 else "f64",
        "data": "synthetic (seeded island-chain Harris field; "
                + ("uniform x, y, mu, log-uniform p)" if focused else "uniform injection)"),
        "config": cfg,
        "roofline": roofline_block(workload, kernel, launch_s, main_leg["launches"], launch_bytes),
    }
    out["roofline"]["bytes_per_step"] = per_step
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        oracle_seed = 7
        out["cpu_baseline"], sample, want = cpu_baseline(xe, ye, f0, f1, model, s, start, seed=oracle_seed)
        log("accuracy of the measured path on the baseline's sample ...")
        out["accuracy"] = accuracy_gate(xe, ye, f0, f1, model, s, nets.get(args.mlp_activation), sample, want,
                                        oracle_seed)
        cfg["matched_spectrum"] = out["accuracy"]["passed"]
    print(json.dumps(out), flush=True)


PARKER_LOOP_PER_FRAME = 1600 * 16  # particles per frame times rank scale
PARKER_LOOP_CAPACITY = 16_000_000


def parker_loop_stored_particles(n, s, frame, seed=160):
    """The stored particles at the start of `frame`: n particles in the state the
    driver injects them in (src/main.cpp:995-1048), uniform in the domain, p = p0,
    weight 1, at the frame's start time."""
    rng = np.random.default_rng(seed)
    parts = kpt.empty_particles(n)
    parts["x"][:] = rng.uniform(s.x_lower, s.x_upper, n)
    parts["y"][:] = rng.uniform(s.y_lower, s.y_upper, n)
    parts["p"][:] = s.p0
    parts["weight"][:] = 1.0
    parts["t"][:] = frame * s.mhd_dt
    parts["step_dt"][:] = s.dt_min
    parts["id"][:] = np.arange(n, dtype=np.uint64)
    parts["e0"][:] = np.sqrt(1.0 + s.p0 * s.p0) - 1.0
    return parts


def run_parker_loop(args):
    """The production Parker loop: a window of frames of the frame loop
    starting from a late-run particle count (see the module docstring)."""
    prec = args.precision or "fp64"
    fp32 = prec == "fp32"
    f0 = args.loop_start_frame
    n0 = f0 * PARKER_LOOP_PER_FRAME
    stretched = args.grid == "stretched"
    y_edges = W.stretched_edges(args.ny, ratio=args.stretch_ratio) if stretched else None
    nframes = args.warmup + args.steps + 1
    fields = [W.island_field(args.nx, args.ny, seed=7, y_edges=y_edges, frame=f0 + k) for k in range(nframes)]
    xe, ye = fields[0][0], fields[0][1]
    h = kpt.default_host_settings()
    h.split_ratio, h.pmin_split_over_p0 = 1.2, 2.0
    s = kpt.make_device_settings(h, xe, ye)
    s.precision = kpt.PREC_FP32 if fp32 else kpt.PREC_FP64
    s.push_chunks = args.chunks
    log(f"bench: Parker loop {args.nx}x{args.ny} {args.grid}, frames {f0}..{f0 + nframes - 2} from {n0} stored "
        f"particles, {PARKER_LOOP_PER_FRAME} injected per frame, {prec}")
    with kpt.Context(PARKER_LOOP_CAPACITY, seed=114514) as ctx:
        ctx.field_upload(kpt.SLOT_LOWER, kpt.MODEL_PARKER16, *fields[0])
        ctx.field_upload(kpt.SLOT_UPPER, kpt.MODEL_PARKER16, *fields[1])
        ctx.upload(parker_loop_stored_particles(n0, s, f0))
        start = ctx.download()
        totals = {"sub_steps": 0, "failures": 0, "split": 0, "skipped": 0, "injected": 0, "histograms": 0}
        kernel_ms, launches, active_in = 0.0, 0, []
        t_begin = None
        for k in range(nframes - 1):
            f = f0 + k
            if k == args.warmup:
                t_begin = time.perf_counter()
            if k > 0:
                ctx.field_advance()
                ctx.field_upload(kpt.SLOT_UPPER, kpt.MODEL_PARKER16, *fields[k + 1])
            t0 = f * s.mhd_dt
            inj = ctx.inject(s, PARKER_LOOP_PER_FRAME, t0, f)
            st = ctx.push_interval(s, t0, f)
            split, skipped = ctx.split(s)
            if (f + 1) % 10 == 0:
                ctx.histogram_logp(128, h.pmin, h.pmax)
            log(f"  frame {f}: {st.particle_steps} sub-steps, {st.kernel_ms:.1f} ms push, stored {ctx.count}, "
                f"split {split}")
            if k >= args.warmup:
                totals["sub_steps"] += st.particle_steps
                totals["failures"] += st.failures
                totals["split"] += split
                totals["skipped"] += skipped
                totals["injected"] += inj
                totals["histograms"] += int((f + 1) % 10 == 0)
                kernel_ms += st.kernel_ms
                launches += st.launches
                active_in.append(st.active_in)
        elapsed = time.perf_counter() - t_begin
        stored = ctx.count
    value = totals["sub_steps"] / elapsed
    launch_s = kernel_ms / 1e3 / launches
    per_step = step_bytes(kpt.MODEL_PARKER16, fp32, stretched=stretched)
    launch_bytes = totals["sub_steps"] / launches * per_step + float(np.mean(active_in)) * \
        state_bytes(kpt.MODEL_PARKER16)
    grid_tag = "_stretched" if stretched else ""
    workload = f"parker_loop_{prec}_{args.nx}x{args.ny}{grid_tag}_island_from_{n0}"
    out = {
        "metric": f"particle-steps/s, production Parker loop, {prec.upper()}",
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
        "data": "synthetic (seeded drifting island-chain frame series; uniform injection)",
        "config": {
            "workload": workload,
            "field": f"{args.nx}x{args.ny}, " + (f"y faces tanh-stretched toward y = 0 (spacing ratio "
                                                 f"{args.stretch_ratio:g})" if stretched else "uniform grid")
                     + ", 2 time-blended frames, the upper frame uploaded every step",
            "frames": f"{f0 + args.warmup}..{f0 + nframes - 2} timed (one bench step = one frame: field advance, "
                      f"injection, push, split, histogram every 10 frames)",
            "stored_at_start": n0,
            "stored_at_end": int(stored),
            "injected_per_frame": PARKER_LOOP_PER_FRAME,
            "capacity": PARKER_LOOP_CAPACITY,
            "split_ratio": h.split_ratio,
            "pmin_split_over_p0": h.pmin_split_over_p0,
            "push_chunks": args.chunks,
            **{k: int(v) for k, v in totals.items()},
        },
        "roofline": roofline_block(workload, "push_kernel", launch_s, launches, launch_bytes),
    }
    out["roofline"]["bytes_per_step"] = per_step
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        # one interval of the stored particles, from the window's first frame pair
        sample = {k: v.copy() for k, v in start.items()}
        sample["t"][:] = 0.0
        xe1, ye1, p0 = fields[0]
        _, _, p1 = fields[1]
        out["cpu_baseline"], smp, want = cpu_baseline(xe1, ye1, p0, p1, kpt.MODEL_PARKER16, s, sample, seed=7)
        log("accuracy of the measured path on the baseline's sample ...")
        out["accuracy"] = accuracy_gate(xe1, ye1, p0, p1, kpt.MODEL_PARKER16, s, None, smp, want, 7)
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
    ap.add_argument("--loop-start-frame", type=int, default=160,
                    help="parker: first frame of the timed window (stored particles = frames x 25 600)")
    ap.add_argument("--parker-push-only", action="store_true",
                    help="parker: push 4 194 304 particles injected once, without the frame loop")
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
    ap.add_argument("--mlp-mu-lo", type=float, default=0.0,
                    help="lower |mu| of the network's training hull (0: the library default)")
    ap.add_argument("--mlp-activation", choices=sorted(kpt.MLP_ACTIVATIONS), default="silu",
                    help="hidden-layer activation of the measured learned closure (the other one is "
                         "pushed as a further leg)")
    ap.add_argument("--no-alt-activation-leg", action="store_true",
                    help="focused with --closure mlp: skip the leg with the other activation")
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
    if args.workload == "parker" and not args.parker_push_only:
        run_parker_loop(args)
        return
    run_push(args)


if __name__ == "__main__":
    main()
