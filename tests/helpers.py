"""Shared helpers of the test suite (fields, particle sets, statistics)."""
from __future__ import annotations

import numpy as np

import kpt

SQRT3 = np.sqrt(3.0)


def uniform_edges(n, lo, hi):
    return np.linspace(lo, hi, n + 1)


def const_field(nx, ny, bx=1.0, by=0.0, bz=0.0, vx=0.0, vy=0.0, vz=0.0):
    shape = (ny, nx)
    return [np.full(shape, v, dtype=np.float64) for v in (bx, by, bz, vx, vy, vz)]


def settings_for(host: kpt.HostSettings, xe, ye) -> kpt.DeviceSettings:
    return kpt.make_device_settings(host, xe, ye)


def random_particles(n, s: kpt.DeviceSettings, rng, p_range=(0.5, 2.0), t0=0.0, mu=True):
    parts = kpt.empty_particles(n)
    parts["x"][:] = rng.uniform(s.x_lower, s.x_upper, n)
    parts["y"][:] = rng.uniform(s.y_lower, s.y_upper, n)
    parts["p"][:] = np.exp(rng.uniform(np.log(p_range[0]), np.log(p_range[1]), n))
    parts["t"][:] = t0
    parts["step_dt"][:] = s.dt_min
    parts["weight"][:] = 1.0
    parts["mu"][:] = rng.uniform(-0.9, 0.9, n) if mu else 0.0
    parts["status"][:] = 0
    parts["split_level"][:] = 0
    parts["id"][:] = np.arange(n, dtype=np.uint64)
    return parts


def copy_parts(parts):
    return {k: v.copy() for k, v in parts.items()}


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)


from accuracy import chi2_two_sample  # noqa: E402,F401  (shared with bench.py)
