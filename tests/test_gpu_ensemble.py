"""The batched ensemble (kpt_ensemble_*): every run of one persistent launch
equals the same run pushed alone through kpt_push_interval (Parker FP32,
focused FP32 with the exact slab-QLT closure and with the learned closure; runs
of uneven sizes, an empty run), bit for bit; a hipGraph-replayed frame loop
equals the eager loop after every replay; the autotuner leaves the particles as
they were and no launch shape changes a result."""
import numpy as np
import pytest

import kpt
import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "p", "t", "step_dt", "weight", "mu", "status", "split_level", "id")
COUNTS = [1000, 0, 2500, 777, 300, 1500]
KINDS = ["parker", "parker", "qlt", "qlt", "mlp", "mlp"]
FRAME = 3


def _hull(act=kpt.MLP_SILU):
    h = kpt.MlpHull()
    h.activation = act
    h.mu_lo, h.log_xi_lo, h.log_xi_hi = 0.02, -4.0, 4.0
    h.gamma_lo, h.gamma_hi, h.h0 = 1.5, 2.0, 1.0e-6
    h.out_mean, h.out_scale = -3.0, 1.0
    return h


def _params():
    return np.random.default_rng(42).normal(0.0, 0.15, kpt.MLP_PARAMS).astype(np.float32)


def _specs(nx=48, ny=24):
    specs = W.ensemble_runs(seed=9, runs=len(COUNTS), nx=nx, ny=ny)
    for spec, n, kind in zip(specs, COUNTS, KINDS):
        spec["count"] = n
        spec["focused"] = kind != "parker"
        spec["kind"] = kind
    return specs


def _settings(spec):
    closure = {"parker": None, "qlt": kpt.CLOSURE_QLT, "mlp": kpt.CLOSURE_MLP}[spec["kind"]]
    return W.ensemble_run_settings(spec, steps_per_interval=20, closure=closure)


def _build(c, specs):
    e = kpt.Ensemble(c, [max(s["count"], 1) for s in specs])
    for r, spec in enumerate(specs):
        s = _settings(spec)
        model = kpt.MODEL_FOCUSED21 if spec["focused"] else kpt.MODEL_PARKER16
        e.field_upload(r, model, spec["xe"], spec["ye"], spec["planes"])
        e.set_run(r, s, spec["seed"])
        e.upload(r, W.ensemble_run_particles(spec, s, frame=FRAME))
    return e


def _alone(spec, act=kpt.MLP_SILU):
    s = _settings(spec)
    parts = W.ensemble_run_particles(spec, s, frame=FRAME)
    model = kpt.MODEL_FOCUSED21 if spec["focused"] else kpt.MODEL_PARKER16
    with kpt.Context(max(spec["count"], 1), seed=spec["seed"]) as c:
        if spec["kind"] == "mlp":
            c.mlp_set(_params(), _hull(act))
        c.field_upload(kpt.SLOT_LOWER, model, spec["xe"], spec["ye"], spec["planes"])
        c.upload(parts)
        st = c.push_interval(s, FRAME * s.mhd_dt, FRAME)
        return c.download(), st


@pytest.mark.parametrize("act", [kpt.MLP_SILU, kpt.MLP_HARD_SWISH])
def test_ensemble_runs_equal_standalone_pushes(act):
    specs = _specs()
    with kpt.Context(16) as c:
        c.mlp_set(_params(), _hull(act))
        with _build(c, specs) as e:
            stats, ms = e.push(FRAME, frames=1)
            got = [e.download(r) for r in range(len(specs))]
    assert ms > 0.0
    for r, spec in enumerate(specs):
        want, st = _alone(spec, act)
        assert stats[r]["particle_steps"] == st.particle_steps, r
        assert stats[r]["failures"] == st.failures, r
        assert stats[r]["active_in"] == st.active_in, r
        for k in FIELDS:
            np.testing.assert_array_equal(got[r][k], want[k], err_msg=f"run {r} ({spec['kind']}) {k}")
    assert sum(s["particle_steps"] for s in stats) > 0


def test_graph_replays_equal_eager_loop():
    specs = _specs()
    with kpt.Context(16) as c:
        c.mlp_set(_params(), _hull())
        with _build(c, specs) as ea, _build(c, specs) as eb:
            for k in range(3):
                sa, _ = ea.push(FRAME + k, frames=1, graph=False)
                sb, _ = eb.push(FRAME + k, frames=1, graph=True)
                assert sa == sb, k
                for r in range(len(specs)):
                    a, b = ea.download(r), eb.download(r)
                    for f in FIELDS:
                        np.testing.assert_array_equal(a[f], b[f], err_msg=f"replay {k} run {r} {f}")
            # three replays in one call equal three eager frames
            sa, _ = ea.push(FRAME + 3, frames=3, graph=False)
            sb, _ = eb.push(FRAME + 3, frames=3, graph=True)
            assert sa == sb
            for r in range(len(specs)):
                a, b = ea.download(r), eb.download(r)
                for f in FIELDS:
                    np.testing.assert_array_equal(a[f], b[f], err_msg=f"3 frames run {r} {f}")
                if specs[r]["count"]:
                    np.testing.assert_allclose(a["t"][a["status"] == 0], (FRAME + 6) * 0.1, rtol=1e-12)


def test_autotune_restores_state_and_shapes_do_not_change_results():
    specs = _specs()
    with kpt.Context(16) as c:
        c.mlp_set(_params(), _hull())
        with _build(c, specs) as e:
            before = [e.download(r) for r in range(len(specs))]
            chosen, table = e.autotune(FRAME, window=0.5)
            assert len(table) == 8 and all(t["ms"] > 0.0 for t in table)
            assert chosen in table and chosen["ms"] == min(t["ms"] for t in table)
            for r in range(len(specs)):
                now = e.download(r)
                for f in FIELDS:
                    np.testing.assert_array_equal(now[f], before[r][f], err_msg=f"run {r} {f}")
            results = []
            for grid, groups in ((0, 8), (97, 1), (1024, 8)):
                for r, spec in enumerate(specs):
                    e.upload(r, before[r])
                e.set_tuning(grid, groups)
                st, _ = e.push(FRAME, frames=1, graph=grid == 97)
                results.append((st, [e.download(r) for r in range(len(specs))]))
    for st, parts in results[1:]:
        assert st == results[0][0]
        for r in range(len(specs)):
            for f in FIELDS:
                np.testing.assert_array_equal(parts[r][f], results[0][1][r][f])


def test_ensemble_histogram_and_rejected_runs():
    specs = _specs()
    with kpt.Context(16) as c:
        c.mlp_set(_params(), _hull())
        with _build(c, specs) as e:
            e.push(FRAME)
            parts = e.download(2)
            h = e.histogram_logp(2, 32, 1.0, 1.0e3)
            act = (parts["status"] == 0) & (parts["p"] >= 1.0) & (parts["p"] <= 1.0e3)
            assert h.sum() == pytest.approx(parts["weight"][act].sum(), rel=1e-12)
            bad = _settings(specs[0])
            bad.precision = kpt.PREC_FP64
            with pytest.raises(kpt.KptError):
                e.set_run(0, bad, 1)
            red = _settings(specs[2])
            red.closure = kpt.CLOSURE_REDUCED
            e.set_run(2, red, 1)
            with pytest.raises(kpt.KptError):
                e.push(FRAME)


def test_sorted_ensemble_equals_standalone_by_id():
    """The per-run cell sort only reorders: after sort + push every particle id
    carries the same state as in the unsorted standalone push."""
    specs = _specs()
    with kpt.Context(16) as c:
        c.mlp_set(_params(), _hull())
        with _build(c, specs) as e:
            act = e.sort()
            assert act == sum(s["count"] for s in specs)
            e.push(FRAME, frames=1)
            got = [e.download(r) for r in range(len(specs))]
    for r, spec in enumerate(specs):
        want, _ = _alone(spec)
        if spec["count"] == 0:
            continue
        og, ow = np.argsort(got[r]["id"]), np.argsort(want["id"])
        assert not np.array_equal(got[r]["id"], want["id"]) or spec["count"] < 300
        for k in FIELDS:
            np.testing.assert_array_equal(got[r][k][og], want[k][ow], err_msg=f"run {r} ({spec['kind']}) {k}")
