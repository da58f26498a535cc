"""ctypes binding of the CPU oracle (oracle/kpt_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker or the timed CPU
baseline. The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "kokkos-particletransport_amd"))
import kpt  # noqa: E402  (struct layouts of include/kpt.h only)

_lib = None


def _has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in f" {line.strip()} " and " avx512vl " in f" {line.strip()} "
    except OSError:
        pass
    return False


def lib_path() -> str:
    v4 = os.path.join(HERE, "libkpt_oracle_v4.so")
    if _has_avx512() and os.path.exists(v4):
        return v4
    return os.path.join(HERE, "libkpt_oracle.so")


def load():
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise OSError(f"{path} is missing: build it with `make -C oracle`")
        lib = C.CDLL(path)
        P = C.POINTER
        dp = P(C.c_double)
        sig = {
            "kpto_philox4x32_10": (None, [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]),
            "kpto_uniform_sqrt3": (C.c_double, [C.c_uint32]),
            "kpto_field_create": (C.c_void_p, [C.c_int32, C.c_int32] + [dp] * 8 + [C.c_int32]),
            "kpto_field_destroy": (None, [C.c_void_p]),
            "kpto_field_coefficients": (C.c_int64, [C.c_void_p, dp, C.c_size_t]),
            "kpto_field_sample": (None, [C.c_void_p, C.c_double, C.c_double, dp]),
            "kpto_push_interval": (None, [C.c_void_p, C.c_void_p, P(kpt.DeviceSettings), C.c_double,
                                          C.c_uint32, C.c_uint64, P(kpt.ParticlesHost), C.c_uint64,
                                          dp, C.c_uint32, C.c_int, P(C.c_uint64), P(C.c_uint64)]),
            "kpto_inject": (C.c_uint64, [P(kpt.DeviceSettings), P(kpt.InjectOpts), C.c_uint64,
                                         C.c_double, C.c_uint32, C.c_uint64, C.c_uint64,
                                         P(kpt.ParticlesHost), C.c_uint64, C.c_uint64]),
            "kpto_split": (C.c_uint64, [P(kpt.DeviceSettings), P(kpt.ParticlesHost), C.c_uint64,
                                        C.c_uint64, P(C.c_uint64), P(C.c_uint64)]),
            "kpto_split_child_id": (C.c_uint64, [C.c_uint64, C.c_int32]),
            "kpto_histogram_logp": (None, [P(kpt.ParticlesHost), C.c_uint64, C.c_int32, C.c_double,
                                           C.c_double, dp]),
            "kpto_histogram_xy": (None, [P(kpt.ParticlesHost), C.c_uint64, C.c_int32, C.c_int32,
                                         C.c_double, C.c_double, C.c_double, C.c_double, dp]),
            "kpto_summary": (None, [P(kpt.ParticlesHost), C.c_uint64, P(kpt.Summary)]),
            "kpto_qlt_scattering": (None, [C.c_double] * 7 + [dp]),
            "kpto_reflect_mu": (C.c_double, [C.c_double, C.c_double]),
            "kpto_split_energy_growth": (C.c_uint64, [P(kpt.ParticleProperties), C.c_double, C.c_double,
                                                      P(kpt.ParticlesHost), C.c_uint64, C.c_uint64, C.c_uint64,
                                                      P(C.c_uint64)]),
            "kpto_reset_split_baselines": (None, [P(kpt.ParticleProperties), P(kpt.ParticlesHost), C.c_uint64]),
            "kpto_apply_boundaries": (C.c_uint64, [C.c_int32, C.c_int32, dp, dp, C.c_int32,
                                                   P(kpt.ParticlesHost), C.c_uint64]),
            "kpto_candidates_field": (C.c_uint64, [dp, C.c_int32, C.c_int32, dp, dp, dp, C.c_int32, C.c_int32,
                                                   C.c_double, P(C.c_uint32), dp, dp]),
            "kpto_candidates_grid": (None, [C.c_int32, C.c_int32, dp, dp, dp, dp]),
            "kpto_generator_inject_candidates": (None, [C.c_int32, P(C.c_uint32), dp, C.c_uint64, C.c_double,
                                                        C.c_int32, C.c_int32, dp, dp, C.c_uint64,
                                                        P(kpt.InjectionKinematics), C.c_double, C.c_uint64,
                                                        C.c_uint32, C.c_uint64, P(kpt.ParticlesHost), C.c_uint64]),
            "kpto_generator_inject_uniform": (C.c_uint64, [dp, C.c_int32, C.c_int32, dp, dp, dp, C.c_int32, C.c_int32,
                                                           C.c_int32, C.c_double, C.c_int32, C.c_uint64,
                                                           P(kpt.InjectionKinematics), C.c_double, C.c_uint64,
                                                           C.c_uint32, C.c_uint64, P(kpt.ParticlesHost),
                                                           C.c_uint64]),
            "kpto_generator_recycle": (C.c_uint64, [C.c_int32, C.c_int32, P(C.c_uint32), dp, C.c_uint64, C.c_double,
                                                    dp, C.c_int32, C.c_int32, dp, dp, dp, C.c_int32, C.c_int32,
                                                    C.c_int32, C.c_double, C.c_int32, C.c_int32,
                                                    P(kpt.InjectionKinematics), C.c_double, C.c_uint64, C.c_uint32,
                                                    P(kpt.ParticlesHost), C.c_uint64, P(C.c_uint64)]),
            "kpto_generic_default_params": (None, [P(kpt.GenericParkerParams)]),
            "kpto_generic_focus_default_params": (None, [P(kpt.GenericFocusParams)]),
            "kpto_generic_focus_create": (C.c_void_p, [C.c_int32, C.c_int32, dp, dp, P(dp), dp,
                                                       P(kpt.GenericFocusParams)]),
            "kpto_generic_focus_time_step": (C.c_double, [C.c_void_p, P(kpt.GenericFocusParams),
                                                          P(kpt.ParticlesHost), C.c_uint64]),
            "kpto_generic_focus_terms": (C.c_int, [C.c_void_p, P(kpt.GenericFocusParams), C.c_double, C.c_double,
                                                   C.c_double, C.c_double, dp]),
            "kpto_generic_focus_step": (C.c_double, [C.c_void_p, P(kpt.GenericFocusParams), P(kpt.ParticlesHost),
                                                     C.c_uint64, C.c_uint32, C.c_uint64, dp, C.c_uint32,
                                                     P(C.c_uint64)]),
            "kpto_generic_create": (C.c_void_p, [C.c_int32, C.c_int32, dp, dp, P(dp),
                                                 P(kpt.GenericParkerParams)]),
            "kpto_generic_time_step": (C.c_double, [C.c_void_p, P(kpt.GenericParkerParams),
                                                    P(kpt.ParticlesHost), C.c_uint64]),
            "kpto_generic_step": (C.c_double, [C.c_void_p, P(kpt.GenericParkerParams),
                                               P(kpt.ParticlesHost), C.c_uint64, C.c_uint32,
                                               C.c_uint64, dp, C.c_uint32, P(C.c_uint64)]),
            "kpto_max_threads": (C.c_int32, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    load().kpto_philox4x32_10(c, k, o)
    return list(o)


def uniform_sqrt3(r: int) -> float:
    return load().kpto_uniform_sqrt3(r)


def qlt_scattering(mu, omega, speed, sigma2, lc, gk=5.0 / 3.0, h0=1e-6):
    """Slab-QLT (D_mumu, dD_mumu/dmu) of the oracle."""
    out = np.zeros(2)
    load().kpto_qlt_scattering(mu, omega, speed, sigma2, lc, gk, h0, _dp(out))
    return out[0], out[1]


def reflect_mu(mu, upper=1.0):
    return load().kpto_reflect_mu(mu, upper)


class Field:
    """One packed coefficient frame on the ghosted grid."""

    def __init__(self, x_edges, y_edges, planes, model):
        lib = load()
        self.xe = np.ascontiguousarray(x_edges, dtype=np.float64)
        self.ye = np.ascontiguousarray(y_edges, dtype=np.float64)
        ps = [np.ascontiguousarray(a, dtype=np.float64) for a in planes]
        self.nx, self.ny, self.model = len(self.xe) - 1, len(self.ye) - 1, model
        self.h = lib.kpto_field_create(self.nx, self.ny, _dp(self.xe), _dp(self.ye),
                                       *[_dp(a) for a in ps], model)
        if not self.h:
            raise ValueError("oracle field creation failed")

    def __del__(self):
        if getattr(self, "h", None):
            load().kpto_field_destroy(self.h)
            self.h = None

    def coefficients(self):
        out = np.empty((self.ny + 2) * (self.nx + 2) * self.model)
        n = load().kpto_field_coefficients(self.h, _dp(out), out.size)
        assert n == out.size
        return out.reshape(self.ny + 2, self.nx + 2, self.model)

    def sample(self, qx, qy):
        out = np.zeros(32)
        load().kpto_field_sample(self.h, qx, qy, _dp(out))
        return out[: self.model]


def push_interval(lower: Field, upper: Field | None, s, t0, frame, seed, parts, noise=None,
                  threads=None):
    """Advance parts (dict of numpy arrays, modified in place). Returns (steps, failures)."""
    n = len(parts["x"])
    h = kpt.particles_struct(parts)
    steps, fails = C.c_uint64(), C.c_uint64()
    nz = None
    nsteps = 0
    if noise is not None:
        nz = np.ascontiguousarray(noise, dtype=np.float64)
        assert nz.ndim == 3 and nz.shape[0] >= n and nz.shape[2] == 4
        nsteps = nz.shape[1]
        assert s.max_particle_steps_per_interval <= nsteps
    threads = threads or os.cpu_count() or 1
    load().kpto_push_interval(lower.h, upper.h if upper else None, C.byref(s), t0, frame, seed,
                              C.byref(h), n, _dp(nz), nsteps, threads, C.byref(steps),
                              C.byref(fails))
    return steps.value, fails.value


def inject(s, opts, requested, t0, frame, seed, first_root, parts, count, capacity):
    h = kpt.particles_struct(parts)
    return load().kpto_inject(C.byref(s), C.byref(opts) if opts else None, requested, t0, frame,
                              seed, first_root, C.byref(h), count, capacity)


def split_child_id(pid, level):
    return load().kpto_split_child_id(pid, level)


def split(s, parts, count, capacity):
    h = kpt.particles_struct(parts)
    a, b = C.c_uint64(), C.c_uint64()
    n = load().kpto_split(C.byref(s), C.byref(h), count, capacity, C.byref(a), C.byref(b))
    return n, a.value, b.value


_SPLIT_STATS = ("active_count", "split_count", "skipped_capacity_count", "skipped_invalid_baseline_count",
                "skipped_minimum_weight_count")


def split_energy_growth(props, policy, parts, count, capacity, first_root):
    """ParticleSystem::split_particles_on_energy_growth on a capacity-sized dict.
    Returns (new count, statistics dict)."""
    h = kpt.particles_struct(parts)
    st = (C.c_uint64 * 5)()
    n = load().kpto_split_energy_growth(C.byref(props), policy.energy_split_ratio, policy.minimum_child_weight,
                                        C.byref(h), count, capacity, first_root, st)
    return n, dict(zip(_SPLIT_STATS, list(st)))


def reset_split_baselines(props, parts, count):
    load().kpto_reset_split_baselines(C.byref(props), C.byref(kpt.particles_struct(parts)), count)


def apply_boundaries(x_edges, y_edges, mask, parts, count):
    xe = np.ascontiguousarray(x_edges, dtype=np.float64)
    ye = np.ascontiguousarray(y_edges, dtype=np.float64)
    return load().kpto_apply_boundaries(len(xe) - 1, len(ye) - 1, _dp(xe), _dp(ye), mask,
                                        C.byref(kpt.particles_struct(parts)), count)


# ---- generic particle generator (include/ParticleGenerator.hpp) ---------------
def _raw(planes):
    return np.ascontiguousarray(np.stack([np.asarray(p, dtype=np.float64) for p in planes]))


def _u32p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


def _region(region):
    return None if region is None else np.ascontiguousarray(region, dtype=np.float64)


def candidates_field(planes, x_edges, y_edges, pred, region=None):
    """(cells, cumulative weights, total) of the field-threshold candidate set;
    pred = (component, comparison, threshold)."""
    raw = _raw(planes)
    xe = np.ascontiguousarray(x_edges, dtype=np.float64)
    ye = np.ascontiguousarray(y_edges, dtype=np.float64)
    n = (len(xe) - 1) * (len(ye) - 1)
    cells, cum, tot = np.empty(n, dtype=np.uint32), np.empty(n), C.c_double()
    reg = _region(region)
    m = load().kpto_candidates_field(_dp(raw), len(xe) - 1, len(ye) - 1, _dp(xe), _dp(ye),
                                     None if reg is None else _dp(reg), pred[0], pred[1], pred[2], _u32p(cells),
                                     _dp(cum), C.byref(tot))
    return cells[:m].copy(), cum[:m].copy(), tot.value


def candidates_grid(x_edges, y_edges):
    """(cumulative cell areas, total) over every cell, x fastest."""
    xe = np.ascontiguousarray(x_edges, dtype=np.float64)
    ye = np.ascontiguousarray(y_edges, dtype=np.float64)
    cum, tot = np.empty((len(xe) - 1) * (len(ye) - 1)), C.c_double()
    load().kpto_candidates_grid(len(xe) - 1, len(ye) - 1, _dp(xe), _dp(ye), _dp(cum), C.byref(tot))
    return cum, tot.value


def generator_inject_candidates(kind, cells, cum, total, x_edges, y_edges, n, kin, e0, seed, call, first_root,
                                parts, first):
    xe = np.ascontiguousarray(x_edges, dtype=np.float64)
    ye = np.ascontiguousarray(y_edges, dtype=np.float64)
    cum = np.ascontiguousarray(cum, dtype=np.float64)
    cells = None if cells is None else np.ascontiguousarray(cells, dtype=np.uint32)
    load().kpto_generator_inject_candidates(kind, _u32p(cells), _dp(cum), len(cum), total, len(xe) - 1,
                                            len(ye) - 1, _dp(xe), _dp(ye), n, C.byref(kin), e0, seed, call,
                                            first_root, C.byref(kpt.particles_struct(parts)), first)


def generator_inject_uniform(planes, x_edges, y_edges, region, pred, max_attempts, n, kin, e0, seed, call,
                             first_root, parts, first):
    raw = _raw(planes)
    xe = np.ascontiguousarray(x_edges, dtype=np.float64)
    ye = np.ascontiguousarray(y_edges, dtype=np.float64)
    reg = _region(region)
    p = pred or (0, 0, 0.0)
    return load().kpto_generator_inject_uniform(_dp(raw), len(xe) - 1, len(ye) - 1, _dp(xe), _dp(ye),
                                                None if reg is None else _dp(reg), int(pred is not None), p[0], p[1],
                                                p[2], max_attempts, n, C.byref(kin), e0, seed, call, first_root,
                                                C.byref(kpt.particles_struct(parts)), first)


def generator_recycle(mode, kin, e0, seed, call, parts, stored, selection, x_edges, y_edges, kind=0, cells=None,
                      cum=None, total=0.0, planes=None, region=None, pred=None, max_attempts=64):
    """(recycled, rejected); mode 0 from the candidate set, 1 uniform with the predicate."""
    xe = np.ascontiguousarray(x_edges, dtype=np.float64)
    ye = np.ascontiguousarray(y_edges, dtype=np.float64)
    raw = _raw(planes) if planes is not None else np.zeros(1)
    cum = np.ascontiguousarray(cum if cum is not None else np.zeros(1), dtype=np.float64)
    cells = None if cells is None else np.ascontiguousarray(cells, dtype=np.uint32)
    reg = _region(region)
    p = pred or (0, 0, 0.0)
    rej = C.c_uint64()
    rec = load().kpto_generator_recycle(mode, kind, _u32p(cells), _dp(cum), len(cum) if cum is not None else 0,
                                        total, _dp(raw), len(xe) - 1, len(ye) - 1, _dp(xe), _dp(ye),
                                        None if reg is None else _dp(reg), int(pred is not None), p[0], p[1], p[2],
                                        max_attempts, selection, C.byref(kin), e0, seed, call,
                                        C.byref(kpt.particles_struct(parts)), stored, C.byref(rej))
    return rec, rej.value


def histogram_logp(parts, count, bins, pmin, pmax):
    out = np.empty(bins)
    load().kpto_histogram_logp(C.byref(kpt.particles_struct(parts)), count, bins, pmin, pmax,
                               _dp(out))
    return out


def histogram_xy(parts, count, bx, by, x_lo, x_hi, y_lo, y_hi):
    out = np.empty(bx * by)
    load().kpto_histogram_xy(C.byref(kpt.particles_struct(parts)), count, bx, by, x_lo, x_hi,
                             y_lo, y_hi, _dp(out))
    return out.reshape(by, bx)


def summary(parts, count) -> dict:
    s = kpt.Summary()
    load().kpto_summary(C.byref(kpt.particles_struct(parts)), count, C.byref(s))
    return s.as_dict()


def generic_default_params():
    p = kpt.GenericParkerParams()
    load().kpto_generic_default_params(C.byref(p))
    return p


class GenericField(Field):
    """Packed 18-component basis of the generic Parker solver on a periodic grid
    (make_legacy_transport_coefficient_fields + build_packed_coefficient_basis)."""

    def __init__(self, x_edges, y_edges, planes, params):
        lib = load()
        self.xe = np.ascontiguousarray(x_edges, dtype=np.float64)
        self.ye = np.ascontiguousarray(y_edges, dtype=np.float64)
        self._planes = [np.ascontiguousarray(a, dtype=np.float64) for a in planes]
        self.nx, self.ny, self.model = len(self.xe) - 1, len(self.ye) - 1, 18
        arr = (C.POINTER(C.c_double) * 6)(*[_dp(a) for a in self._planes])
        self.params = params
        self.h = lib.kpto_generic_create(self.nx, self.ny, _dp(self.xe), _dp(self.ye), arr,
                                         C.byref(params))
        if not self.h:
            raise ValueError("generic Parker coefficients: the reference aborts on this field")

    def time_step(self, parts):
        h = kpt.particles_struct(parts)
        return load().kpto_generic_time_step(self.h, C.byref(self.params), C.byref(h),
                                             len(parts["x"]))

    def step(self, parts, step, seed=0, noise=None):
        """One advance_stochastic step (modifies parts). Returns (dt, failures)."""
        h = kpt.particles_struct(parts)
        fails = C.c_uint64()
        nz, ns = None, 0
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            assert nz.ndim == 3 and nz.shape[2] == 4 and nz.shape[0] >= len(parts["x"])
            ns = nz.shape[1]
        dt = load().kpto_generic_step(self.h, C.byref(self.params), C.byref(h), len(parts["x"]),
                                      step, seed, _dp(nz), ns, C.byref(fails))
        return dt, fails.value


def generic_focus_default_params():
    p = kpt.GenericFocusParams()
    load().kpto_generic_focus_default_params(C.byref(p))
    return p


FOCUS_TERMS = ("dt_local", "u_x", "u_y", "p_rate", "mu_rate", "D", "dD", "kpar", "kperp", "keff_x", "keff_y",
               "b_x", "b_y", "b_z")


class GenericFocusField(Field):
    """Packed 27-component basis of the generic focused-transport solver on a
    periodic grid (make_legacy_transport_coefficient_fields_with_time_derivative
    + FocusTransportSolver::build_packed_coefficient_basis)."""

    def __init__(self, x_edges, y_edges, planes, params, dvdt=None):
        lib = load()
        self.xe = np.ascontiguousarray(x_edges, dtype=np.float64)
        self.ye = np.ascontiguousarray(y_edges, dtype=np.float64)
        self._planes = [np.ascontiguousarray(a, dtype=np.float64) for a in planes]
        self.nx, self.ny, self.model = len(self.xe) - 1, len(self.ye) - 1, 27
        self._dvdt = None if dvdt is None else np.ascontiguousarray(np.stack([np.asarray(a, np.float64)
                                                                               for a in dvdt]))
        arr = (C.POINTER(C.c_double) * 6)(*[_dp(a) for a in self._planes])
        self.params = params
        self.h = lib.kpto_generic_focus_create(self.nx, self.ny, _dp(self.xe), _dp(self.ye), arr,
                                               _dp(self._dvdt), C.byref(params))
        if not self.h:
            raise ValueError("generic focused coefficients: the reference aborts on this field")

    def time_step(self, parts):
        h = kpt.particles_struct(parts)
        return load().kpto_generic_focus_time_step(self.h, C.byref(self.params), C.byref(h), len(parts["x"]))

    def terms(self, x, y, p, mu):
        """Local terms of one particle (dict over FOCUS_TERMS), None where the reference aborts."""
        out = np.empty(len(FOCUS_TERMS))
        ok = load().kpto_generic_focus_terms(self.h, C.byref(self.params), x, y, p, mu, _dp(out))
        return dict(zip(FOCUS_TERMS, out.tolist())) if ok else None

    def step(self, parts, step, seed=0, noise=None):
        """One advance_stochastic step (modifies parts). Returns (dt, failures)."""
        h = kpt.particles_struct(parts)
        fails = C.c_uint64()
        nz, ns = None, 0
        if noise is not None:
            nz = np.ascontiguousarray(noise, dtype=np.float64)
            assert nz.ndim == 3 and nz.shape[2] == 4 and nz.shape[0] >= len(parts["x"])
            ns = nz.shape[1]
        dt = load().kpto_generic_focus_step(self.h, C.byref(self.params), C.byref(h), len(parts["x"]), step, seed,
                                            _dp(nz), ns, C.byref(fails))
        return dt, fails.value


def max_threads() -> int:
    """OpenMP threads of the generic solver loops (OMP_NUM_THREADS or every core)."""
    return int(load().kpto_max_threads())
