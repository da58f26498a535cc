"""The kernels the bench times, on the bench's own field, against the FP64 oracle.

Field: the 1024 x 512 island-chain Harris sheet with y faces tanh-stretched toward
y = 0 (spacing ratio 10) and two time-blended frames, as the bench builds it.
Physics: focused transport with slab-QLT scattering (gamma_k 5/3, sigma2 0.1,
Lc 0.02), mu ~ U[-1, 1], p log-uniform on [1, 1e3], fixed dt 5e-5 over one 0.1
interval (2001 sub-steps), pushed in 16 cell-sorted chunks as the bench pushes it;
and Parker transport with the driver's default physics and adaptive dt.

Kernels covered (stretched lookup, STR = true):
  - push_kernel<21, true, float, QLT>: the FP32 exact-QLT push;
  - push_kernel_mlp<STR, ACT, PASSES>: the learned closure, both activations, one
    network evaluation per sub-step (the default) and one per Picard pass;
  - push_kernel<16, false, float>: the FP32 Parker push.

Bar (FP32 is statistical against the FP64 oracle): with independent seeds, the
pitch-angle, log-p and spatial histograms pass a two-sample chi-square with
p > 0.01 and KS on mu, p and y has p > 0.01; since both runs start from one set,
the displacements over the interval (dx, dy, d ln p, d mu), which depend only on
the dynamics, carry the sharper test: KS and chi-square on each. The level is
0.01 for the family of tests of one comparison (Bonferroni: every p-value above
0.01 / the number of tests). With the oracle's seed, the median particle follows the oracle's
trajectory (the FP32 Philox draws differ from the FP64 ones by < 2^-23 sqrt(3)),
which bounds the error per observable; for the learned closure the bound on the
pitch-angle displacement is set by the network's fit (about 1% rms in D)."""
import numpy as np
import pytest
from scipy import stats

import accuracy as A
import kpt
import kpt_oracle as O
import workloads as W
from helpers import chi2_two_sample

pytestmark = pytest.mark.gpu

N_FOCUSED = 65536
N_PARKER = 32768
ORACLE_SEED = 99
OTHER_SEED = 5
P_MIN = 0.01


@pytest.fixture(scope="module")
def field():
    ye = W.stretched_edges(512, ratio=10.0)
    xe, ye, f0 = W.island_field(1024, 512, seed=7, y_edges=ye)
    _, _, f1 = W.island_field(1024, 512, seed=8, y_edges=ye)
    return xe, ye, f0, f1


@pytest.fixture(scope="module")
def focused_oracle(field):
    """The learned-closure workload's physics on the bench field: the start set and the FP64 oracle's interval."""
    xe, ye, f0, f1 = field
    s = W.focused_qlt_device_settings(xe, ye)
    start = W.focused_qlt_particles(N_FOCUSED, s, seed=11)
    want = {k: v.copy() for k, v in start.items()}
    steps, fails = O.push_interval(O.Field(xe, ye, f0, 21), O.Field(xe, ye, f1, 21), s, 0.0, 0, ORACLE_SEED, want)
    assert fails == 0
    return s, start, want, steps


@pytest.fixture(scope="module", params=[kpt.MLP_SILU, kpt.MLP_HARD_SWISH], ids=["silu", "hard-swish"])
def network(request):
    o = kpt.mlp_default_train_opts()
    o.samples = 1 << 20
    o.steps = 4000
    o.hull.activation = request.param
    with kpt.Context(16) as c:
        rep = c.mlp_train(o)
        params, hull = c.mlp_get()
    assert rep["val_rms_rel_d"] < 0.05, rep
    return params, hull


def _gpu(field, model, s, start, seed, net=None):
    xe, ye, f0, f1 = field
    with kpt.Context(len(start["x"]), seed=seed) as c:
        if net is not None:
            c.mlp_set(*net)
        c.field_upload(kpt.SLOT_LOWER, model, xe, ye, f0)
        c.field_upload(kpt.SLOT_UPPER, model, xe, ye, f1)
        c.upload(start)
        st = c.push_interval(s, 0.0, 0)
        got = c.download()
    return A.sort_by_id(got), st  # the chunked push leaves the set cell-sorted


def _pathwise(a, b, s, keys):
    return A.pathwise_medians(a, b, s.x_width, s.y_width, keys)


def _check_focused(field, s, start, want, oracle_steps, net=None):
    same, st = _gpu(field, 21, s, start, ORACLE_SEED, net)
    assert st.failures == 0 and st.launches == 16
    assert st.particle_steps == oracle_steps
    assert np.all(np.abs(same["mu"]) <= 1.0)
    err = _pathwise(same, want, s, ("x", "y", "p", "mu"))
    dmu = float(np.median(np.abs(want["mu"] - start["mu"])))
    print(f"pathwise medians {err}, median |mu - mu_start| {dmu:.3e}")
    other, _ = _gpu(field, 21, s, start, OTHER_SEED, net)
    p = A.focused_pvalues(want, other, start, s)
    print("independent seeds:", p)
    assert min(p.values()) > P_MIN / len(p), p
    return err, dmu


def test_fp32_exact_qlt_stretched_matches_oracle(field, focused_oracle):
    """push_kernel<21, true, float, QLT, STR>: the median particle agrees with the
    FP64 oracle to float rounding (x, y and ln p to 1e-7, the pitch-angle
    displacement to 1e-3 of itself); the statistics match."""
    s, start, want, steps = focused_oracle
    s32 = s.copy()
    s32.precision = kpt.PREC_FP32
    s32.push_chunks = 16
    err, dmu = _check_focused(field, s32, start, want, steps)
    assert err["x"] < 1e-7 and err["y"] < 1e-7 and err["p"] < 1e-7, err
    assert err["mu"] < 1e-3 * dmu, (err, dmu)


@pytest.mark.parametrize("passes", [1, 3])
def test_learned_closure_stretched_matches_oracle(field, focused_oracle, network, passes):
    """push_kernel_mlp<STR, ACT, PASSES> (the bench's headline kernel with
    PASSES = 1): spatial, momentum and pitch-angle statistics match the FP64
    oracle; with the oracle's seed the pitch-angle displacement differs from
    the oracle's by a few percent of itself (the network's error in D enters
    the noise amplitude as its square root), positions and p to float
    rounding of the exact push."""
    s, start, want, steps = focused_oracle
    s32 = s.copy()
    s32.precision = kpt.PREC_FP32
    s32.push_chunks = 16
    s32.closure = kpt.CLOSURE_MLP
    s32.mlp_passes = passes
    err, dmu = _check_focused(field, s32, start, want, steps, net=network)
    assert err["mu"] < 0.05 * dmu, (err, dmu)
    assert err["x"] < 1e-5 and err["y"] < 1e-5 and err["p"] < 1e-6, err


def test_one_network_evaluation_follows_three(field, network):
    """The default closure (one network evaluation per sub-step, the Picard passes
    extended to first order in ln|mu| with the network's slope output) against
    the closure that evaluates the network at every pass, on one injected noise
    table: the pitch-angle increments agree to 1e-3 of their rms (median)."""
    xe, ye, f0, f1 = field
    s = W.focused_qlt_device_settings(xe, ye, closure=kpt.CLOSURE_MLP, precision=kpt.PREC_FP32)
    n, steps = 16384, 8
    s.mhd_dt = steps * s.dt_min
    s.max_particle_steps_per_interval = steps + 1
    start = W.focused_qlt_particles(n, s, seed=3)
    noise = np.random.default_rng(4).uniform(-np.sqrt(3), np.sqrt(3), (n, steps + 1, 4))
    out = {}
    for passes in (1, 3):
        s2 = s.copy()
        s2.mlp_passes = passes
        with kpt.Context(n, seed=1) as c:
            c.mlp_set(*network)
            c.field_upload(kpt.SLOT_LOWER, 21, xe, ye, f0)
            c.field_upload(kpt.SLOT_UPPER, 21, xe, ye, f1)
            c.upload(start)
            c.set_noise(noise)
            st = c.push_interval(s2, 0.0, 0)
            assert st.failures == 0
            out[passes] = c.download()
    d1 = out[1]["mu"] - start["mu"]
    d3 = out[3]["mu"] - start["mu"]
    rel = np.abs(d1 - d3) / np.sqrt(np.mean(d3 ** 2))
    print(f"median / 99th percentile difference of the mu increments over their rms: "
          f"{np.median(rel):.2e} / {np.quantile(rel, 0.99):.2e}")
    assert np.median(rel) < 1e-3 and np.quantile(rel, 0.99) < 2e-2
    np.testing.assert_array_equal(out[1]["t"], out[3]["t"])
    # positions follow mu through the streaming term only
    np.testing.assert_allclose(out[1]["x"], out[3]["x"], rtol=0, atol=1e-6)


def test_fp32_parker_stretched_matches_oracle(field):
    """push_kernel<16, false, float, STR>: Parker transport with the driver's
    default physics and adaptive dt (the local cell width enters the dt limiter)
    on the bench field, injected uniformly: spatial density and log-p spectrum
    chi-square and KS p > 0.01 against the FP64 oracle with independent seeds;
    with the oracle's seed the median particle agrees to 1e-6."""
    xe, ye, f0, f1 = field
    h = kpt.default_host_settings()
    h.split_particles = 0
    s = kpt.make_device_settings(h, xe, ye)
    o = kpt.InjectOpts()
    o.fixed_time = 1
    start = kpt.empty_particles(N_PARKER)
    O.inject(s, o, N_PARKER, 0.0, 0, 114514, 0, start, 0, N_PARKER)
    want = {k: v.copy() for k, v in start.items()}
    steps, fails = O.push_interval(O.Field(xe, ye, f0, 16), O.Field(xe, ye, f1, 16), s, 0.0, 0, ORACLE_SEED, want)
    assert fails == 0
    s32 = s.copy()
    s32.precision = kpt.PREC_FP32
    s32.push_chunks = 16
    same, st = _gpu(field, 16, s32, start, ORACLE_SEED)
    assert st.failures == 0
    assert abs(st.particle_steps / steps - 1.0) < 1e-3, (st.particle_steps, steps)
    err = _pathwise(same, want, s, ("x", "y", "p"))
    print(f"pathwise medians {err}; sub-steps GPU {st.particle_steps} oracle {steps}")
    assert err["x"] < 1e-6 and err["y"] < 1e-6 and err["p"] < 1e-6, err
    other, _ = _gpu(field, 16, s32, start, OTHER_SEED)
    hw = np.histogram2d(want["y"], want["x"], bins=(32, 64), range=((s.y_lower, s.y_upper), (s.x_lower, s.x_upper)))[0]
    hg = np.histogram2d(other["y"], other["x"], bins=(32, 64), range=((s.y_lower, s.y_upper), (s.x_lower, s.x_upper)))[0]
    lp = (np.log(0.99 * want["p"].min()), np.log(1.01 * want["p"].max()))
    pw = np.histogram(np.log(want["p"]), bins=64, range=lp)[0]
    pg = np.histogram(np.log(other["p"]), bins=64, range=lp)[0]
    p = {"chi2_xy": chi2_two_sample(hw, hg)[2], "chi2_logp": chi2_two_sample(pw, pg)[2],
         "ks_y": float(stats.ks_2samp(want["y"], other["y"]).pvalue),
         "ks_p": float(stats.ks_2samp(want["p"], other["p"]).pvalue)}
    p.update(A.displacement_pvalues(want, other, start, s.x_width, s.y_width, ("dx", "dy", "dlnp")))
    print("independent seeds:", p)
    assert min(p.values()) > P_MIN / len(p), p
