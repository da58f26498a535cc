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


def island_field(nx=1024, ny=512, x_range=(-1.0, 1.0), y_range=(-0.5, 0.5), width=0.02,
                 by_amp=0.05, bz=0.1, inflow=0.1, noise=0.01, seed=7, modes=4):
    """Harris sheet with an island chain By = a sum_k sin(k pi x + phi_k)/k exp(-(y/w)^2)
    and 1% Gaussian velocity noise (seeded)."""
    rng = np.random.default_rng(seed)
    phases = rng.uniform(0.0, 2.0 * np.pi, size=modes)
    xe = np.linspace(x_range[0], x_range[1], nx + 1)
    ye = np.linspace(y_range[0], y_range[1], ny + 1)
    x = xe[0] + np.arange(nx) * (xe[-1] - xe[0]) / nx
    y = ye[0] + np.arange(ny) * (ye[-1] - ye[0]) / ny
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


def c1_host_settings() -> "kpt.HostSettings":
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


C1 = dict(nx=256, ny=128, particles=65536, region_y=0.1, hist_bins=64, hist_pmin=0.1,
          hist_pmax=1.0e3, xy_bins=(64, 32))


def c1_inject_opts() -> "kpt.InjectOpts":
    o = kpt.InjectOpts()
    o.use_region = 1
    o.fixed_time = 1
    o.region_x_lo, o.region_x_hi = -1.0, 1.0
    o.region_y_lo, o.region_y_hi = -C1["region_y"], C1["region_y"]
    o.p_init = 1.0
    return o
