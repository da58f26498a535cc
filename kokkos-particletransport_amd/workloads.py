"""Synthetic, seeded workloads of the Parker / focused push (numpy only).

Fields are returned in the compact-field layout (include/AthenaFieldReader.hpp:
274-288): face coordinates x_edges[nx+1], y_edges[ny+1] and six planes
Bx, By, Bz, Vx, Vy, Vz of shape (ny, nx), x fastest. Values are evaluated at
the node positions x0 + i*dx, y0 + j*dy, which is where the driver's sampler
puts value (i, j) (the +dx/2 shift of src/main.cpp:616-617, :1142-1144).
"""
from __future__ import annotations

import numpy as np

import kpt


def harris_field(nx=256, ny=128, x_range=(-1.0, 1.0), y_range=(-0.5, 0.5), width=0.05,
                 by_amp=0.05, bz=0.1, inflow=0.1):
    """Harris current sheet with a sin(pi x) reconnection perturbation.

    Bx = tanh(y/w), By = a sin(pi x) exp(-(y/w)^2), Bz = const,
    Vx = 0, Vy = -u tanh(y/w), Vz = 0.
    """
    xe = np.linspace(x_range[0], x_range[1], nx + 1)
    ye = np.linspace(y_range[0], y_range[1], ny + 1)
    x = xe[0] + np.arange(nx) * (xe[-1] - xe[0]) / nx
    y = ye[0] + np.arange(ny) * (ye[-1] - ye[0]) / ny
    X, Y = np.meshgrid(x, y)  # (ny, nx)
    Bx = np.tanh(Y / width)
    By = by_amp * np.sin(np.pi * X) * np.exp(-((Y / width) ** 2))
    Bz = np.full_like(X, bz)
    Vx = np.zeros_like(X)
    Vy = -inflow * np.tanh(Y / width)
    Vz = np.zeros_like(X)
    return xe, ye, [Bx, By, Bz, Vx, Vy, Vz]


def stretched_edges(n, lo=-0.5, hi=0.5, ratio=10.0):
    """n + 1 faces refined toward the middle of [lo, hi]: y = mid + half atanh(a s) /
    atanh(a) with s uniform on [-1, 1] and a^2 = 1 - 1/ratio, so the spacing grows
    smoothly from the middle to the ends by `ratio` (the grid coordinate s is a
    tanh of y)."""
    a = np.sqrt(1.0 - 1.0 / ratio)
    s = np.linspace(-1.0, 1.0, n + 1)
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    e = mid + half * np.arctanh(a * s) / np.arctanh(a)
    e[0], e[-1] = lo, hi
    return e


def island_field(nx=1024, ny=512, x_range=(-1.0, 1.0), y_range=(-0.5, 0.5), width=0.02,
                 by_amp=0.05, bz=0.1, inflow=0.1, noise=0.01, seed=7, modes=4, y_edges=None, frame=0):
    """Harris sheet with an island chain By = a sum_k sin(k pi x + phi_k)/k exp(-(y/w)^2)
    and 1% Gaussian velocity noise (seeded). Values sit at the nodes x0 + i dx (and,
    with stretched `y_edges`, at the y faces). `frame` > 0 gives a frame of a time
    series: the island phases drift by 0.02 k per frame and the noise is drawn
    afresh."""
    rng = np.random.default_rng(seed)
    phases = rng.uniform(0.0, 2.0 * np.pi, size=modes)
    if frame:
        phases = phases + 0.02 * np.arange(1, modes + 1) * frame
        rng = np.random.default_rng([seed, frame])
    xe = np.linspace(x_range[0], x_range[1], nx + 1)
    ye = np.linspace(y_range[0], y_range[1], ny + 1) if y_edges is None else np.asarray(y_edges, dtype=np.float64)
    x = xe[0] + np.arange(nx) * (xe[-1] - xe[0]) / nx
    y = ye[0] + np.arange(ny) * (ye[-1] - ye[0]) / ny if y_edges is None else ye[:-1]
    X, Y = np.meshgrid(x, y)
    Bx = np.tanh(Y / width)
    By = np.zeros_like(X)
    for k in range(1, modes + 1):
        By += np.sin(k * np.pi * X + phases[k - 1]) / k
    By *= by_amp * np.exp(-((Y / width) ** 2))
    Bz = np.full_like(X, bz)
    Vx = inflow * noise * rng.standard_normal(X.shape)
    Vy = -inflow * np.tanh(Y / width) * (1.0 + noise * rng.standard_normal(X.shape))
    Vz = np.zeros_like(X)
    return xe, ye, [Bx, By, Bz, Vx, Vy, Vz]


def parker_harris_host_settings() -> "kpt.HostSettings":
    """Parker transport, plain FP64, small: kappa_par = 0.01 p^(4/3), kperp/kpar = 0.05,
    fixed dt = 1e-4 over one 0.1 interval (1000 steps), no splitting."""
    h = kpt.default_host_settings()
    h.kpara0 = 0.01
    h.p0 = 1.0
    h.gamma_turb = 5.0 / 3.0
    h.magnetic_dependency = 0
    h.momentum_dependency = 1
    h.kperp_over_kpara = 0.05
    h.mhd_dt = 0.1
    h.dt_min_rel = 1.0e-3
    h.dt_max_rel = 1.0e-3
    h.split_particles = 0
    h.time_interpolation = 1
    h.pmin = 0.1
    h.pmax = 1.0e3
    return h


PARKER_HARRIS_256X128 = dict(nx=256, ny=128, particles=65536, region_y=0.1, hist_bins=64, hist_pmin=0.1,
          hist_pmax=1.0e3, xy_bins=(64, 32))


def parker_harris_inject_opts() -> "kpt.InjectOpts":
    o = kpt.InjectOpts()
    o.use_region = 1
    o.fixed_time = 1
    o.region_x_lo, o.region_x_hi = -1.0, 1.0
    o.region_y_lo, o.region_y_hi = -PARKER_HARRIS_256X128["region_y"], PARKER_HARRIS_256X128["region_y"]
    o.p_init = 1.0
    return o


def focused_qlt_host_settings() -> "kpt.HostSettings":
    """Focused transport on the 1024 x 512 island field: reflecting |mu| <= 1,
    fixed dt = 5e-5 over one 0.1 interval (2000 sub-steps), no splitting.
    The focused equation carries parallel transport explicitly (v mu b plus
    pitch-angle scattering), so only kappa_perp enters; it is held constant at
    kpara0 * kperp/kpar = 0.01 * 0.05 = 5e-4: the field-line random walk that
    sets kappa_perp scales with v, which is within 30% of c over p in [1, 1e3]
    (the small Harris-sheet Parker scaling p^(4/3) would give kappa_perp = 5 at
    p = 1e3)."""
    h = parker_harris_host_settings()
    h.focused_transport = 1
    h.momentum_dependency = 0
    h.dt_min_rel = 5.0e-4
    h.dt_max_rel = 5.0e-4
    h.mu_max = 1.0
    h.pmin = 1.0
    h.pmax = 1.0e3
    return h


FOCUSED_QLT_1024X512 = dict(nx=1024, ny=512, particles=4 * 1024 * 1024, p_range=(1.0, 1.0e3), qlt_sigma2=0.1,
          qlt_correlation_length=0.02, qlt_spectral_index=5.0 / 3.0, qlt_h0=1.0e-6,
          hist_bins=64, mu_bins=32)
FOCUSED_QLT_1024X512_8M = dict(FOCUSED_QLT_1024X512, particles=8 * 1024 * 1024)


def focused_qlt_device_settings(xe, ye, closure=None, precision=None) -> "kpt.DeviceSettings":
    """Slab-QLT scattering: gamma_k = 5/3, sigma2 = 0.1, Lc = 0.02, h0 = 1e-6
    (PitchAngleScatteringParameters, include/LegencyModel.hpp:196-198)."""
    s = kpt.make_device_settings(focused_qlt_host_settings(), xe, ye)
    s.closure = kpt.CLOSURE_QLT if closure is None else closure
    s.precision = kpt.PREC_FP64 if precision is None else precision
    s.qlt_sigma2 = FOCUSED_QLT_1024X512["qlt_sigma2"]
    s.qlt_correlation_length = FOCUSED_QLT_1024X512["qlt_correlation_length"]
    s.qlt_spectral_index = FOCUSED_QLT_1024X512["qlt_spectral_index"]
    s.qlt_h0 = FOCUSED_QLT_1024X512["qlt_h0"]
    return s


def focused_qlt_particles(n, s, seed=11, p_range=None) -> dict:
    """(x, y) uniform on the domain, mu ~ U[-1, 1], p log-uniform on p_range,
    t = 0, weight 1, ids 0..n-1."""
    lo, hi = p_range or FOCUSED_QLT_1024X512["p_range"]
    rng = np.random.default_rng(seed)
    parts = kpt.empty_particles(n)
    parts["x"][:] = rng.uniform(s.x_lower, s.x_upper, n)
    parts["y"][:] = rng.uniform(s.y_lower, s.y_upper, n)
    parts["p"][:] = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    parts["mu"][:] = rng.uniform(-1.0, 1.0, n)
    parts["t"][:] = 0.0
    parts["step_dt"][:] = s.dt_min
    parts["weight"][:] = 1.0
    parts["status"][:] = 0
    parts["split_level"][:] = 0
    parts["id"][:] = np.arange(n, dtype=np.uint64)
    return parts


ENSEMBLE_64X512X256 = dict(runs=64, nx=512, ny=256, count_range=(1.0e3, 1.0e6), bz_range=(0.0, 0.5), ratio_range=(0.01, 0.2),
          inflow_range=(0.05, 0.15), steps_per_interval=50, frames=10)


def ensemble_runs(seed=2024, runs=None, nx=None, ny=None, count_range=None) -> list:
    """The batched-ensemble workload: per run a seeded island-chain field variant
    (guide field Bz ~ U[0, 0.5], inflow ~ U[0.05, 0.15]), kperp/kpar ~ U[0.01, 0.2],
    a particle count log-uniform on count_range; even runs Parker, odd runs
    focused transport with the learned closure."""
    runs = ENSEMBLE_64X512X256["runs"] if runs is None else runs
    nx = ENSEMBLE_64X512X256["nx"] if nx is None else nx
    ny = ENSEMBLE_64X512X256["ny"] if ny is None else ny
    lo, hi = ENSEMBLE_64X512X256["count_range"] if count_range is None else count_range
    rng = np.random.default_rng(seed)
    out = []
    for r in range(runs):
        bz = rng.uniform(*ENSEMBLE_64X512X256["bz_range"])
        ratio = rng.uniform(*ENSEMBLE_64X512X256["ratio_range"])
        inflow = rng.uniform(*ENSEMBLE_64X512X256["inflow_range"])
        count = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        xe, ye, planes = island_field(nx, ny, bz=bz, inflow=inflow, seed=seed * 1000 + r)
        out.append(dict(run=r, bz=bz, kperp_over_kpara=ratio, inflow=inflow, count=count, focused=r % 2 == 1,
                        xe=xe, ye=ye, planes=planes, seed=seed * 1000 + r))
    return out


def ensemble_run_settings(spec, steps_per_interval=None, closure=None) -> "kpt.DeviceSettings":
    """FP32 settings of one ensemble run: the small Harris-sheet Parker physics or the focused
    QLT physics with the run's kperp/kpar, fixed dt = mhd_dt / steps_per_interval,
    one frame (no time interpolation)."""
    n = ENSEMBLE_64X512X256["steps_per_interval"] if steps_per_interval is None else steps_per_interval
    h = focused_qlt_host_settings() if spec["focused"] else parker_harris_host_settings()
    h.kperp_over_kpara = spec["kperp_over_kpara"]
    h.dt_min_rel = h.dt_max_rel = 1.0 / n
    h.time_interpolation = 0
    if spec["focused"]:
        s = focused_qlt_device_settings(spec["xe"], spec["ye"], closure=kpt.CLOSURE_MLP if closure is None else closure,
                               precision=kpt.PREC_FP32)
        s2 = kpt.make_device_settings(h, spec["xe"], spec["ye"])
        for k in ("kperp_over_kpara", "dt_min", "dt_max", "time_interpolation", "kpara0"):
            if hasattr(s, k):
                setattr(s, k, getattr(s2, k))
        return s
    s = kpt.make_device_settings(h, spec["xe"], spec["ye"])
    s.precision = kpt.PREC_FP32
    return s


def ensemble_run_particles(spec, s, frame=0) -> dict:
    """The run's particles at the start of interval `frame`: Parker runs at p = 1
    in |y| < 0.1, focused runs with the focused slab-QLT distribution."""
    n = spec["count"]
    if spec["focused"]:
        parts = focused_qlt_particles(n, s, seed=spec["seed"])
    else:
        rng = np.random.default_rng(spec["seed"])
        parts = kpt.empty_particles(n)
        parts["x"][:] = rng.uniform(s.x_lower, s.x_upper, n)
        parts["y"][:] = rng.uniform(-0.1, 0.1, n)
        parts["p"][:] = 1.0
        parts["step_dt"][:] = s.dt_min
        parts["weight"][:] = 1.0
        parts["id"][:] = np.arange(n, dtype=np.uint64)
    parts["t"][:] = frame * s.mhd_dt
    return parts
