"""The batched ensemble at the 64-run workload's shape, against the FP64 oracle.

Workload: 64 independent runs, each on its own seeded 512 x 256 island-chain field
variant (guide field, inflow and kperp/kpar drawn per run), particle counts
log-uniform on [1e3, 1e6], even runs Parker and odd runs focused transport with
the learned closure, FP32, a 500-sub-step frame loop (10 intervals of 50 fixed
steps) with a cell sort of every run before each interval and the interval
replayed from a hipGraph, as the bench runs it.

1. Each run kind (Parker FP32, focused FP32 with the exact slab-QLT closure,
   focused FP32 with a network trained on the GPU) on such field variants through
   that loop, against FP64 oracle runs of the same particles: with the run's own
   seed the median particle follows the oracle's trajectory; with another seed
   the displacement statistics match (KS and chi-square, family-wise level 0.01).
2. The full 64-run shape: every run's sub-step count equals its particle count times
   the sub-steps one particle takes over the loop (the interval arithmetic
   replayed in double), no failures, every particle at the loop's end time,
   each run's ids a permutation of its start ids, and each run's momentum
   histogram total equal to its active in-range weight."""
import numpy as np
import pytest

import accuracy as A
import kpt
import kpt_oracle as O
import workloads as W

pytestmark = pytest.mark.gpu

FRAMES = W.ENSEMBLE_64X512X256["frames"]
STEPS = W.ENSEMBLE_64X512X256["steps_per_interval"]


@pytest.fixture(scope="module")
def network():
    o = kpt.mlp_default_train_opts()
    o.samples = 1 << 20
    o.steps = 3000
    o.hull.activation = kpt.MLP_SILU
    with kpt.Context(16) as c:
        rep = c.mlp_train(o)
        net = c.mlp_get()
    assert rep["val_rms_rel_d"] < 0.05, rep
    return net


def _loop(e, frames=FRAMES):
    """The bench's frame loop: cell sort of every run, then one interval replayed
    from the graph. Returns the per-run stats summed over the intervals."""
    tot = None
    for k in range(frames):
        e.sort()
        st, _ = e.push(k, frames=1, graph=True)
        tot = st if tot is None else [{f: a[f] + b[f] for f in a} for a, b in zip(tot, st)]
    return tot


def _oracle(spec, s, start, seed, frames=FRAMES):
    s64 = s.copy()
    s64.precision = kpt.PREC_FP64
    if s64.closure == kpt.CLOSURE_MLP:
        s64.closure = kpt.CLOSURE_QLT
    f = O.Field(spec["xe"], spec["ye"], spec["planes"], 21 if spec["focused"] else 16)
    parts = {k: v.copy() for k, v in start.items()}
    steps = 0
    for k in range(frames):
        n, fails = O.push_interval(f, f, s64, k * s64.mhd_dt, k, seed, parts)
        assert fails == 0
        steps += n
    return parts, steps


def test_ensemble_run_kinds_match_oracle(network):
    specs = W.ensemble_runs(seed=31, runs=3)
    kinds = ["parker", "qlt", "mlp"]
    for spec, kind in zip(specs, kinds):
        spec["count"] = 20000
        spec["focused"] = kind != "parker"
    settings = [W.ensemble_run_settings(spec, closure={"parker": None, "qlt": kpt.CLOSURE_QLT, "mlp": kpt.CLOSURE_MLP}[kind])
                for spec, kind in zip(specs, kinds)]
    starts = [W.ensemble_run_particles(spec, s, frame=0) for spec, s in zip(specs, settings)]
    with kpt.Context(16, seed=1) as c:
        c.mlp_set(*network)
        with kpt.Ensemble(c, [spec["count"] for spec in specs]) as e:
            for r, (spec, s, start) in enumerate(zip(specs, settings, starts)):
                model = kpt.MODEL_FOCUSED21 if spec["focused"] else kpt.MODEL_PARKER16
                e.field_upload(r, model, spec["xe"], spec["ye"], spec["planes"])
                e.set_run(r, s, spec["seed"])
                e.upload(r, start)
            stats = _loop(e)
            got = [A.sort_by_id(e.download(r)) for r in range(len(specs))]
    for r, (spec, kind, s, start) in enumerate(zip(specs, kinds, settings, starts)):
        assert stats[r]["failures"] == 0
        same, steps = _oracle(spec, s, start, spec["seed"])
        other, _ = _oracle(spec, s, start, spec["seed"] + 1000)
        assert stats[r]["particle_steps"] == steps, (kind, stats[r]["particle_steps"], steps)
        err = A.pathwise_medians(got[r], same, s.x_width, s.y_width)
        keys = ("dx", "dy", "dlnp", "dmu") if spec["focused"] else ("dx", "dy", "dlnp")
        p = A.displacement_pvalues(other, got[r], start, s.x_width, s.y_width, keys)
        print(kind, "pathwise medians", err, "p-values", p)
        assert min(p.values()) > 0.01 / len(p), (kind, p)
        assert err["x"] < 1e-5 * s.x_width and err["y"] < 1e-5 * s.y_width and err["p"] < 1e-5, (kind, err)
        if spec["focused"]:
            dmu = float(np.median(np.abs(same["mu"] - start["mu"])))
            assert err["mu"] < (0.05 if kind == "mlp" else 1e-3) * dmu, (kind, err, dmu)


def _sub_steps_per_particle(t_start, mhd_dt, dt, frames):
    """Sub-steps of one particle over the frame loop, replaying the push's
    interval arithmetic in double: intervals [k mhd_dt, k mhd_dt + mhd_dt), fixed
    dt truncated to the interval end."""
    t, n = t_start, 0
    for k in range(frames):
        t0 = k * mhd_dt
        t_end = t0 + mhd_dt
        while t < t_end:
            step = dt if dt <= t_end - t else t_end - t
            t += step
            n += 1
    return n, t


def test_ensemble_64_runs_properties(network):
    specs = W.ensemble_runs()
    settings = [W.ensemble_run_settings(spec) for spec in specs]
    starts = [W.ensemble_run_particles(spec, s, frame=0) for spec, s in zip(specs, settings)]
    total = sum(spec["count"] for spec in specs)
    assert len(specs) == 64 and total > 1_000_000
    with kpt.Context(16, seed=2) as c:
        c.mlp_set(*network)
        with kpt.Ensemble(c, [spec["count"] for spec in specs]) as e:
            for r, (spec, s, start) in enumerate(zip(specs, settings, starts)):
                model = kpt.MODEL_FOCUSED21 if spec["focused"] else kpt.MODEL_PARKER16
                e.field_upload(r, model, spec["xe"], spec["ye"], spec["planes"])
                e.set_run(r, s, spec["seed"])
                e.upload(r, start)
            stats = _loop(e)
            hists = [e.histogram_logp(r, 64, 0.1, 1.0e3) for r in range(len(specs))]
            got = [e.download(r) for r in range(len(specs))]
    for r, (spec, s, start) in enumerate(zip(specs, settings, starts)):
        per, t_end = _sub_steps_per_particle(0.0, s.mhd_dt, s.dt_min, FRAMES)
        assert stats[r]["failures"] == 0, r
        assert stats[r]["particle_steps"] == spec["count"] * per, (r, stats[r]["particle_steps"], per)
        q = got[r]
        assert len(q["x"]) == spec["count"]
        assert np.all(q["status"] == 0), r
        np.testing.assert_array_equal(q["t"], np.full(spec["count"], t_end))
        np.testing.assert_array_equal(np.sort(q["id"]), np.sort(start["id"]))
        in_range = (q["p"] >= 0.1) & (q["p"] <= 1.0e3)
        assert hists[r].sum() == pytest.approx(q["weight"][in_range].sum(), rel=1e-12), r
        if spec["focused"]:
            assert np.all(np.abs(q["mu"]) <= 1.0), r
