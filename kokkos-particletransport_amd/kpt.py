"""ctypes binding of libkpt.so (the C ABI of include/kpt.h) for tests and bench.

The structures mirror include/kpt.h field by field. The library is loaded
lazily; `load()` raises if libkpt.so is missing, so nothing here can fall back
to a CPU path: every compute call goes to the gfx950 kernels.

Particle sets are dicts of numpy arrays with the keys of PARTICLE_FIELDS.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("KPT_LIB") or os.path.join(HERE, "libkpt.so")

KPT_OK = 0
MODEL_PARKER16 = 16
MODEL_FOCUSED21 = 21
SLOT_LOWER = 0
SLOT_UPPER = 1
PREC_FP64 = 0
PREC_FP32 = 1
CLOSURE_REDUCED = 0
CLOSURE_QLT = 1
CLOSURE_MLP = 2
NOALLOC_MAX_CHUNKS = 16          # KPT_NOALLOC_MAX_CHUNKS
NOALLOC_SCRATCH_BYTES = 24576    # KPT_NOALLOC_SCRATCH_BYTES

PARTICLE_FIELDS = (
    ("x", np.float64), ("y", np.float64), ("p", np.float64), ("t", np.float64),
    ("step_dt", np.float64), ("weight", np.float64), ("mu", np.float64),
    ("status", np.int32), ("split_level", np.int32), ("id", np.uint64), ("e0", np.float64),
)
# ParticleSystem arrays (include/ParticleSystem.hpp:342-349): both previous
# positions and the sort key. Optional in a particle dict: absent on upload, the
# previous positions take the current one and the key 0.
SYSTEM_FIELDS = (
    ("prev_x", np.float64), ("prev_y", np.float64), ("prev_step_x", np.float64), ("prev_step_y", np.float64),
    ("sort_key", np.int32),
)


class DeviceSettings(C.Structure):
    """kpt_device_settings: Reconnection2DDeviceSettings (src/main.cpp:246-283)."""
    _fields_ = [
        ("max_particle_steps_per_interval", C.c_int32),
        ("momentum_dependency", C.c_int32),
        ("magnetic_dependency", C.c_int32),
        ("split_particles", C.c_int32),
        ("time_interpolation", C.c_int32),
        ("focused_transport", C.c_int32),
        ("x_lower", C.c_double), ("x_upper", C.c_double),
        ("y_lower", C.c_double), ("y_upper", C.c_double),
        ("x_width", C.c_double), ("y_width", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double),
        ("interpolation_shift_x", C.c_double), ("interpolation_shift_y", C.c_double),
        ("mhd_dt", C.c_double), ("p0", C.c_double), ("pmax", C.c_double),
        ("gamma_turb", C.c_double), ("pindex", C.c_double), ("kpara0", C.c_double),
        ("kperp_over_kpara", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("drift_param1", C.c_double), ("drift_param2", C.c_double), ("charge", C.c_double),
        ("particle_v0", C.c_double), ("duu0", C.c_double),
        ("pitch_angle_scattering_h0", C.c_double), ("split_ratio", C.c_double),
        ("pmin_split", C.c_double), ("mu_max", C.c_double),
        ("precision", C.c_int32), ("push_chunks", C.c_int32),
        ("closure", C.c_int32), ("mlp_passes", C.c_int32),
        ("qlt_sigma2", C.c_double), ("qlt_correlation_length", C.c_double),
        ("qlt_spectral_index", C.c_double), ("qlt_h0", C.c_double),
        ("fp64_null_b", C.c_double),
    ]

    def copy(self) -> "DeviceSettings":
        out = DeviceSettings()
        C.pointer(out)[0] = self
        return out


class HostSettings(C.Structure):
    """kpt_host_settings: physics subset of Reconnection2DSettings (src/main.cpp:114-154)."""
    _fields_ = [
        ("max_particle_steps_per_interval", C.c_int32),
        ("momentum_dependency", C.c_int32),
        ("magnetic_dependency", C.c_int32),
        ("split_particles", C.c_int32),
        ("time_interpolation", C.c_int32),
        ("focused_transport", C.c_int32),
        ("mhd_dt", C.c_double), ("p0", C.c_double), ("pmin", C.c_double), ("pmax", C.c_double),
        ("gamma_turb", C.c_double), ("kpara0", C.c_double), ("kperp_over_kpara", C.c_double),
        ("dt_min_rel", C.c_double), ("dt_max_rel", C.c_double),
        ("drift_param1", C.c_double), ("drift_param2", C.c_double), ("charge", C.c_double),
        ("particle_v0", C.c_double), ("duu0", C.c_double), ("split_ratio", C.c_double),
        ("pmin_split_over_p0", C.c_double), ("mu_max", C.c_double),
    ]


class ParticlesHost(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("p", C.POINTER(C.c_double)), ("t", C.POINTER(C.c_double)),
        ("step_dt", C.POINTER(C.c_double)), ("weight", C.POINTER(C.c_double)),
        ("mu", C.POINTER(C.c_double)), ("status", C.POINTER(C.c_int32)),
        ("split_level", C.POINTER(C.c_int32)), ("id", C.POINTER(C.c_uint64)),
        ("e0", C.POINTER(C.c_double)),
        ("prev_x", C.POINTER(C.c_double)), ("prev_y", C.POINTER(C.c_double)),
        ("prev_step_x", C.POINTER(C.c_double)), ("prev_step_y", C.POINTER(C.c_double)),
        ("sort_key", C.POINTER(C.c_int32)),
    ]


class PushStats(C.Structure):
    _fields_ = [
        ("particle_steps", C.c_uint64), ("failures", C.c_uint64), ("active_in", C.c_uint64),
        ("kernel_ms", C.c_double), ("launches", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Summary(C.Structure):
    """kpt_summary: ParticleSummary (src/main.cpp:338-350)."""
    _fields_ = [
        ("stored", C.c_uint64), ("active", C.c_uint64), ("inactive", C.c_uint64),
        ("weight_sum", C.c_double), ("p_min", C.c_double), ("p_max", C.c_double),
        ("p_avg", C.c_double), ("dt_min", C.c_double), ("dt_max", C.c_double),
        ("dt_avg", C.c_double), ("max_split_level", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class FrameTiming(C.Structure):
    """kpt_frame_timing: FrameTiming (src/main.cpp:355-363)."""
    _fields_ = [
        ("read_seconds", C.c_double), ("inject_seconds", C.c_double),
        ("move_seconds", C.c_double), ("split_seconds", C.c_double),
        ("injected", C.c_uint64), ("split", C.c_uint64), ("skipped_split_capacity", C.c_uint64),
    ]


class InjectOpts(C.Structure):
    _fields_ = [
        ("use_region", C.c_int32), ("fixed_time", C.c_int32),
        ("region_x_lo", C.c_double), ("region_x_hi", C.c_double),
        ("region_y_lo", C.c_double), ("region_y_hi", C.c_double),
        ("p_init", C.c_double),
    ]


class GenericParkerParams(C.Structure):
    """kpt_generic_parker_params: model controls of the generic Parker solver."""
    _fields_ = [
        ("perpendicular_model", C.c_int32), ("reserved", C.c_int32),
        ("parallel_prefactor", C.c_double), ("perpendicular_prefactor", C.c_double),
        ("perpendicular_parallel_ratio", C.c_double),
        ("minimum_magnetic_field_magnitude", C.c_double), ("minimum_solar_wind_speed", C.c_double),
        ("minimum_sigma2", C.c_double), ("minimum_magnetic_field_squared_for_drift", C.c_double),
        ("sigma2", C.c_double), ("l2d", C.c_double), ("lslab", C.c_double), ("lc", C.c_double),
        ("rest_mass", C.c_double), ("charge", C.c_double), ("speed_of_light", C.c_double),
        ("courant", C.c_double),
    ]

    def copy(self) -> "GenericParkerParams":
        c = GenericParkerParams()
        C.pointer(c)[0] = self
        return c


class GenericFocusParams(C.Structure):
    """kpt_generic_focus_params: the generic focused-transport solver's controls."""
    _fields_ = [
        ("model", GenericParkerParams), ("minimum_magnetic_field_magnitude", C.c_double),
        ("spectral_index", C.c_double), ("resonance_regularization_h0", C.c_double),
        ("scattering_minimum_magnetic_field_magnitude", C.c_double), ("scattering_minimum_sigma2", C.c_double),
        ("scattering_minimum_correlation_length", C.c_double), ("scattering_minimum_speed", C.c_double),
        ("maximum_pitch_angle_cosine", C.c_double),
    ]

    def copy(self) -> "GenericFocusParams":
        c = GenericFocusParams()
        C.pointer(c)[0] = self
        return c


class EnsembleTuning(C.Structure):
    """kpt_ensemble_tuning: launch shape of the ensemble's persistent push."""
    _fields_ = [("grid_blocks", C.c_int32), ("groups", C.c_int32), ("ms", C.c_double)]

    def as_dict(self) -> dict:
        return {"grid_blocks": self.grid_blocks, "groups": self.groups, "ms": self.ms}


class EnsembleRunStats(C.Structure):
    _fields_ = [("particle_steps", C.c_uint64), ("failures", C.c_uint64), ("active_in", C.c_uint64)]


class ParticleProperties(C.Structure):
    """kpt_particle_properties (ParticleProperties, include/ParticleSystem.hpp:145-213)."""
    _fields_ = [("rest_mass", C.c_double), ("charge", C.c_double), ("speed_of_light", C.c_double),
                ("energy_scale_erg", C.c_double)]


class EnergySplitPolicy(C.Structure):
    """kpt_energy_split_policy (ParticleSplittingPolicy, include/ParticleSystem.hpp:280-295)."""
    _fields_ = [("energy_split_ratio", C.c_double), ("minimum_child_weight", C.c_double)]


# FieldPredicateComparison and ParticleRecycleSelection (include/ParticleGenerator.hpp:117-133)
CMP_GREATER, CMP_GREATER_EQUAL, CMP_LESS, CMP_LESS_EQUAL, CMP_ABS_GREATER, CMP_ABS_LESS = range(6)
RECYCLE_ESCAPED, RECYCLE_INACTIVE, RECYCLE_ANY = range(3)


class InjectionKinematics(C.Structure):
    """kpt_injection_kinematics (ParticleInjectionKinematics, include/ParticleGenerator.hpp:67-74,
    plus the time and stored step the driver's push reads)."""
    _fields_ = [("momentum", C.c_double), ("mu", C.c_double), ("weight", C.c_double), ("status", C.c_int32),
                ("pad", C.c_int32), ("time", C.c_double), ("step_dt", C.c_double)]


class Region(C.Structure):
    """kpt_region (AxisAlignedParticleRegion, include/ParticleGenerator.hpp:17-60)."""
    _fields_ = [("x_lo", C.c_double), ("x_hi", C.c_double), ("y_lo", C.c_double), ("y_hi", C.c_double)]


class FieldPredicate(C.Structure):
    """kpt_field_predicate (FieldThresholdParticlePredicate, include/ParticleGenerator.hpp:254-358)."""
    _fields_ = [("component", C.c_int32), ("comparison", C.c_int32), ("threshold", C.c_double)]


def injection_kinematics(momentum=1.0, mu=0.0, weight=1.0, status=0, time=0.0, step_dt=0.0) -> InjectionKinematics:
    return InjectionKinematics(momentum, mu, weight, status, 0, time, step_dt)


class SplitStatistics(C.Structure):
    """kpt_split_statistics (ParticleSplitStatistics, include/ParticleSystem.hpp:268-275)."""
    _fields_ = [("active_count", C.c_uint64), ("split_count", C.c_uint64),
                ("skipped_capacity_count", C.c_uint64), ("skipped_invalid_baseline_count", C.c_uint64),
                ("skipped_minimum_weight_count", C.c_uint64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MlpHull(C.Structure):
    """kpt_mlp_hull: training hull, output scaling and activation of the
    pitch-angle network."""
    _fields_ = [
        ("mu_lo", C.c_double), ("log_xi_lo", C.c_double), ("log_xi_hi", C.c_double),
        ("gamma_lo", C.c_double), ("gamma_hi", C.c_double), ("h0", C.c_double),
        ("out_mean", C.c_double), ("out_scale", C.c_double),
        ("activation", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class MlpTrainOpts(C.Structure):
    _fields_ = [
        ("samples", C.c_uint64), ("steps", C.c_uint32), ("batch", C.c_uint32),
        ("lr", C.c_double), ("seed", C.c_uint64), ("hull", MlpHull),
    ]


class MlpReport(C.Structure):
    _fields_ = [
        ("final_loss", C.c_double), ("val_rms_lnr", C.c_double), ("val_max_lnr", C.c_double),
        ("val_rms_y1", C.c_double), ("val_max_y1", C.c_double), ("val_rms_rel_d", C.c_double),
        ("train_ms", C.c_double), ("steps", C.c_uint32), ("validation_samples", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


MLP_PARAMS = 8706
MLP_SILU, MLP_HARD_SWISH = 0, 1
MLP_ACTIVATIONS = {"silu": MLP_SILU, "hard-swish": MLP_HARD_SWISH}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def particles_struct(parts: dict) -> ParticlesHost:
    """ParticlesHost pointing into a dict of contiguous numpy arrays (None -> NULL)."""
    h = ParticlesHost()
    for name, dt in PARTICLE_FIELDS + SYSTEM_FIELDS:
        a = parts.get(name)
        if a is None:
            continue
        assert a.dtype == dt and a.flags.c_contiguous, name
        ctype = {np.float64: C.c_double, np.int32: C.c_int32, np.uint64: C.c_uint64}[dt]
        setattr(h, name, a.ctypes.data_as(C.POINTER(ctype)))
    return h


def empty_particles(n: int, system=False) -> dict:
    """Zeroed particle dict; `system` adds the ParticleSystem arrays."""
    fields = PARTICLE_FIELDS + (SYSTEM_FIELDS if system else ())
    return {name: np.zeros(n, dtype=dt) for name, dt in fields}


# ---------------------------------------------------------------------------
# library loading and signatures
# ---------------------------------------------------------------------------
_lib = None

EXPORTS = {
    "kpt_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_uint64, C.c_uint64]),
    "kpt_destroy": (C.c_int, [C.c_void_p]),
    "kpt_last_error": (C.c_char_p, [C.c_void_p]),
    "kpt_last_create_error": (C.c_char_p, []),
    "kpt_version": (C.c_char_p, []),
    "kpt_device_name": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "kpt_default_host_settings": (C.c_int, [C.POINTER(HostSettings)]),
    "kpt_make_device_settings": (C.c_int, [C.POINTER(HostSettings), C.c_int32, C.c_int32,
                                           C.POINTER(C.c_double), C.POINTER(C.c_double),
                                           C.POINTER(DeviceSettings)]),
    "kpt_set_seed": (C.c_int, [C.c_void_p, C.c_uint64]),
    "kpt_field_load_compact": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32]),
    "kpt_field_load_compact_ex": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32]),
    "kpt_field_upload": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
                         + [C.POINTER(C.c_double)] * 8),
    "kpt_field_advance": (C.c_int, [C.c_void_p]),
    "kpt_field_stage_compact": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32]),
    "kpt_field_advance_staged": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "kpt_field_download_coefficients": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                                  C.c_size_t]),
    "kpt_compact_field_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
                               + [C.POINTER(C.c_double)] * 8 + [C.c_int32, C.c_char_p, C.c_size_t]),
    "kpt_compact_field_write": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32]
                                + [C.POINTER(C.c_double)] * 8),
    "kpt_particles_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "kpt_particles_capacity": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "kpt_particles_upload": (C.c_int, [C.c_void_p, C.POINTER(ParticlesHost), C.c_uint64]),
    "kpt_particles_download": (C.c_int, [C.c_void_p, C.POINTER(ParticlesHost), C.c_uint64]),
    "kpt_inject": (C.c_int, [C.c_void_p, C.POINTER(DeviceSettings), C.POINTER(InjectOpts),
                             C.c_uint64, C.c_double, C.c_uint32, C.POINTER(C.c_uint64)]),
    "kpt_push_interval": (C.c_int, [C.c_void_p, C.POINTER(DeviceSettings), C.c_double,
                                    C.c_uint32, C.POINTER(PushStats)]),
    "kpt_push_interval_noalloc": (C.c_int, [C.c_void_p, C.POINTER(DeviceSettings), C.c_double,
                                            C.c_uint32, C.POINTER(PushStats)]),
    "kpt_split": (C.c_int, [C.c_void_p, C.POINTER(DeviceSettings), C.POINTER(C.c_uint64),
                            C.POINTER(C.c_uint64)]),
    "kpt_histogram_logp": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double,
                                     C.POINTER(C.c_double)]),
    "kpt_histogram_xy": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.POINTER(C.c_double)]),
    "kpt_summary_compute": (C.c_int, [C.c_void_p, C.POINTER(Summary)]),
    "kpt_sort_particles": (C.c_int, [C.c_void_p, C.POINTER(DeviceSettings), C.c_int32,
                                     C.POINTER(C.c_uint64)]),
    "kpt_set_noise": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_uint64, C.c_uint32]),
    "kpt_clear_noise": (C.c_int, [C.c_void_p]),
    "kpt_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]),
    "kpt_snapshot_write": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_double, C.c_double]),
    "kpt_snapshot_write_resumable": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_double, C.c_double]),
    "kpt_snapshot_read": (C.c_int, [C.c_void_p, C.c_char_p, C.c_double, C.c_double, C.c_double,
                                    C.c_double]),
    "kpt_write_histogram_csv": (C.c_int, [C.c_char_p, C.c_int32, C.c_double, C.c_double,
                                          C.POINTER(C.c_double)]),
    "kpt_format_summary_header": (C.c_int, [C.c_char_p, C.c_size_t]),
    "kpt_format_summary_row": (C.c_int, [C.c_char_p, C.c_size_t, C.c_int32, C.c_double,
                                         C.POINTER(Summary), C.POINTER(FrameTiming), C.c_double,
                                         C.c_double]),
    "kpt_snapshot_write_host": (C.c_int, [C.c_char_p, C.POINTER(ParticlesHost), C.c_uint64,
                                          C.c_uint64, C.c_double, C.c_double, C.c_double]),
    "kpt_snapshot_read_host": (C.c_int, [C.c_char_p, C.POINTER(ParticlesHost),
                                         C.POINTER(C.c_uint64), C.c_double, C.c_double,
                                         C.c_char_p, C.c_size_t]),
    "kpt_generic_parker_default_params": (C.c_int, [C.POINTER(GenericParkerParams)]),
    "kpt_generic_parker_build": (C.c_int, [C.c_void_p, C.POINTER(GenericParkerParams), C.c_int32]),
    "kpt_generic_parker_download": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                              C.c_size_t]),
    "kpt_generic_parker_time_step": (C.c_int, [C.c_void_p, C.POINTER(GenericParkerParams), C.c_int32,
                                               C.POINTER(C.c_double)]),
    "kpt_generic_parker_step": (C.c_int, [C.c_void_p, C.POINTER(GenericParkerParams), C.c_int32,
                                          C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "kpt_generic_focus_default_params": (C.c_int, [C.POINTER(GenericFocusParams)]),
    "kpt_generic_focus_build": (C.c_int, [C.c_void_p, C.POINTER(GenericFocusParams), C.c_int32,
                                          C.POINTER(C.c_double)]),
    "kpt_generic_focus_download": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_size_t]),
    "kpt_generic_focus_time_step": (C.c_int, [C.c_void_p, C.POINTER(GenericFocusParams), C.c_int32,
                                              C.POINTER(C.c_double)]),
    "kpt_generic_focus_step": (C.c_int, [C.c_void_p, C.POINTER(GenericFocusParams), C.c_int32, C.c_uint32,
                                         C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "kpt_generic_focus_advance_deterministic": (C.c_int, [C.c_void_p, C.POINTER(GenericFocusParams), C.c_int32,
                                                          C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "kpt_generic_parker_advance_deterministic": (C.c_int, [C.c_void_p, C.POINTER(GenericParkerParams), C.c_int32,
                                                           C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "kpt_capture_previous_positions": (C.c_int, [C.c_void_p]),
    "kpt_sort_by_key": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "kpt_compact_active_particles": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "kpt_particle_system_write_snapshot": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(ParticleProperties),
                                                     C.c_int32, C.POINTER(EnergySplitPolicy)]),
    "kpt_particle_system_read_snapshot": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(ParticleProperties),
                                                    C.POINTER(C.c_int32), C.POINTER(EnergySplitPolicy)]),
    "kpt_ensemble_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_void_p)]),
    "kpt_ensemble_destroy": (C.c_int, [C.c_void_p]),
    "kpt_ensemble_field_upload": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
                                  + [C.POINTER(C.c_double)] * 8),
    "kpt_ensemble_set_run": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(DeviceSettings), C.c_uint64]),
    "kpt_ensemble_particles_upload": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ParticlesHost), C.c_uint64]),
    "kpt_ensemble_particles_download": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(ParticlesHost), C.c_uint64]),
    "kpt_ensemble_count": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]),
    "kpt_ensemble_inject": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.c_double, C.c_uint32,
                                      C.POINTER(C.c_uint64)]),
    "kpt_ensemble_push": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(EnsembleRunStats),
                                    C.POINTER(C.c_double)]),
    "kpt_ensemble_autotune": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.POINTER(EnsembleTuning),
                                        C.POINTER(EnsembleTuning), C.c_int32, C.POINTER(C.c_int32)]),
    "kpt_ensemble_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(EnsembleTuning)]),
    "kpt_ensemble_histogram_logp": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                              C.POINTER(C.c_double)]),
    "kpt_ensemble_sort": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "kpt_default_particle_properties": (C.c_int, [C.POINTER(ParticleProperties)]),
    "kpt_default_energy_split_policy": (C.c_int, [C.POINTER(EnergySplitPolicy)]),
    "kpt_split_energy_growth": (C.c_int, [C.c_void_p, C.POINTER(ParticleProperties),
                                          C.POINTER(EnergySplitPolicy), C.POINTER(SplitStatistics)]),
    "kpt_reset_split_baselines": (C.c_int, [C.c_void_p, C.POINTER(ParticleProperties)]),
    "kpt_apply_boundaries": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_uint64)]),
    "kpt_candidates_from_field": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Region), C.POINTER(FieldPredicate),
                                            C.POINTER(C.c_void_p)]),
    "kpt_candidates_from_grid": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "kpt_candidates_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double)]),
    "kpt_candidates_download": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "kpt_candidates_destroy": (C.c_int, [C.c_void_p]),
    "kpt_inject_from_candidates": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(InjectionKinematics),
                                             C.POINTER(ParticleProperties), C.POINTER(C.c_uint64)]),
    "kpt_inject_uniform_if": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Region), C.POINTER(FieldPredicate), C.c_int32,
                                        C.c_uint64, C.POINTER(InjectionKinematics), C.POINTER(ParticleProperties),
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "kpt_recycle_uniform_if": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Region), C.POINTER(FieldPredicate),
                                         C.c_int32, C.c_int32, C.POINTER(InjectionKinematics),
                                         C.POINTER(ParticleProperties), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64)]),
    "kpt_recycle_from_candidates": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(InjectionKinematics),
                                              C.POINTER(ParticleProperties), C.POINTER(C.c_uint64)]),
    "kpt_generator_calls": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    "kpt_particle_kinematics": (C.c_int, [C.c_void_p, C.POINTER(ParticleProperties), C.POINTER(C.c_double),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_uint64]),
    "kpt_mlp_default_train_opts": (C.c_int, [C.POINTER(MlpTrainOpts)]),
    "kpt_mlp_train": (C.c_int, [C.c_void_p, C.POINTER(MlpTrainOpts), C.POINTER(MlpReport)]),
    "kpt_mlp_set": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_size_t, C.POINTER(MlpHull)]),
    "kpt_mlp_get": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_size_t, C.POINTER(MlpHull)]),
    "kpt_mlp_gradient": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_double)]),
    "kpt_mlp_infer": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_uint64, C.c_int32,
                                C.POINTER(C.c_float)]),
}


def load():
    """Load libkpt.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build it with `make -C kokkos-particletransport_amd`")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class KptError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"kpt error {code}: {msg}")
        self.code = code


class Candidates:
    """A kpt_candidates set (FieldParticleCandidateSet / GridParticleCandidateSet)."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        n, t = C.c_uint64(), C.c_double()
        _check_static(ctx.lib.kpt_candidates_info(h, C.byref(n), C.byref(t)))
        self.count, self.total = n.value, t.value

    def download(self):
        """(cell indices i + nx j, cumulative weights)"""
        cells = np.empty(self.count, dtype=np.uint32)
        cum = np.empty(self.count)
        self.ctx._check(self.ctx.lib.kpt_candidates_download(
            self.ctx.h, self.h, cells.ctypes.data_as(C.POINTER(C.c_uint32)), _dp(cum)))
        return cells, cum

    def close(self):
        if self.h:
            self.ctx.lib.kpt_candidates_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def default_host_settings() -> HostSettings:
    h = HostSettings()
    _check_static(load().kpt_default_host_settings(C.byref(h)))
    return h


def make_device_settings(host: HostSettings, x_edges, y_edges) -> DeviceSettings:
    xe = np.ascontiguousarray(x_edges, dtype=np.float64)
    ye = np.ascontiguousarray(y_edges, dtype=np.float64)
    out = DeviceSettings()
    _check_static(load().kpt_make_device_settings(C.byref(host), len(xe) - 1, len(ye) - 1,
                                                  _dp(xe), _dp(ye), C.byref(out)))
    return out


def generic_default_params() -> GenericParkerParams:
    p = GenericParkerParams()
    _check_static(load().kpt_generic_parker_default_params(C.byref(p)))
    return p


def generic_focus_default_params() -> GenericFocusParams:
    p = GenericFocusParams()
    _check_static(load().kpt_generic_focus_default_params(C.byref(p)))
    return p


def default_particle_properties() -> ParticleProperties:
    p = ParticleProperties()
    _check_static(load().kpt_default_particle_properties(C.byref(p)))
    return p


def default_energy_split_policy() -> EnergySplitPolicy:
    p = EnergySplitPolicy()
    _check_static(load().kpt_default_energy_split_policy(C.byref(p)))
    return p


def mlp_default_train_opts() -> MlpTrainOpts:
    o = MlpTrainOpts()
    _check_static(load().kpt_mlp_default_train_opts(C.byref(o)))
    return o


def philox4x32_10(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    load().kpt_philox4x32_10(c, k, o)
    return list(o)


def compact_field_read(path, allow_nonuniform=False):
    lib = load()
    nx, ny = C.c_int32(), C.c_int32()
    err = C.create_string_buffer(512)
    st = lib.kpt_compact_field_read(path.encode(), C.byref(nx), C.byref(ny), *([None] * 8),
                                    int(allow_nonuniform), err, 512)
    if st != KPT_OK:
        raise KptError(st, err.value.decode())
    n0, n1 = nx.value, ny.value
    xe = np.empty(n0 + 1)
    ye = np.empty(n1 + 1)
    planes = [np.empty((n1, n0)) for _ in range(6)]
    st = lib.kpt_compact_field_read(path.encode(), C.byref(nx), C.byref(ny), _dp(xe), _dp(ye),
                                    *[_dp(a) for a in planes], int(allow_nonuniform), err, 512)
    if st != KPT_OK:
        raise KptError(st, err.value.decode())
    return xe, ye, planes


def compact_field_write(path, x_edges, y_edges, planes):
    xe = np.ascontiguousarray(x_edges, dtype=np.float64)
    ye = np.ascontiguousarray(y_edges, dtype=np.float64)
    ps = [np.ascontiguousarray(a, dtype=np.float64) for a in planes]
    st = load().kpt_compact_field_write(path.encode(), len(xe) - 1, len(ye) - 1, _dp(xe), _dp(ye),
                                        *[_dp(a) for a in ps])
    _check_static(st)


def write_histogram_csv(path, pmin, pmax, weights):
    w = np.ascontiguousarray(weights, dtype=np.float64)
    _check_static(load().kpt_write_histogram_csv(path.encode(), len(w), pmin, pmax, _dp(w)))


def summary_csv_header() -> str:
    buf = C.create_string_buffer(1024)
    _check_static(load().kpt_format_summary_header(buf, 1024))
    return buf.value.decode()


def summary_csv_row(frame, time, summary: Summary, timing: FrameTiming, total_frame_s,
                    total_elapsed_s) -> str:
    buf = C.create_string_buffer(2048)
    _check_static(load().kpt_format_summary_row(buf, 2048, frame, time, C.byref(summary),
                                                C.byref(timing), total_frame_s, total_elapsed_s))
    return buf.value.decode()


def snapshot_write_host(path, parts: dict, count, capacity, charge, split_ratio, p0):
    h = particles_struct(parts)
    _check_static(load().kpt_snapshot_write_host(path.encode(), C.byref(h), count, capacity,
                                                 charge, split_ratio, p0))


def snapshot_read_host(path, charge, split_ratio) -> dict:
    """Read a v4 snapshot into a particle dict (t and step_dt are left zero)."""
    lib = load()
    n = C.c_uint64(0)
    err = C.create_string_buffer(512)
    st = lib.kpt_snapshot_read_host(path.encode(), None, C.byref(n), charge, split_ratio, err, 512)
    if st != KPT_OK:
        raise KptError(st, err.value.decode())
    parts = empty_particles(n.value)
    h = particles_struct(parts)
    st = lib.kpt_snapshot_read_host(path.encode(), C.byref(h), C.byref(n), charge, split_ratio,
                                    err, 512)
    if st != KPT_OK:
        raise KptError(st, err.value.decode())
    return parts


def _check_static(st):
    if st != KPT_OK:
        raise KptError(st, "call failed")


class Context:
    """One kpt_ctx: particle storage, two coefficient frames and the RNG seed."""

    def __init__(self, capacity: int, seed: int = 114514, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.kpt_create(C.byref(h), device, capacity, seed)
        if st != KPT_OK:
            raise KptError(st, self.lib.kpt_last_create_error().decode())
        self.h = h

    def close(self):
        if self.h:
            self.lib.kpt_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != KPT_OK:
            raise KptError(st, self.lib.kpt_last_error(self.h).decode())

    @property
    def count(self) -> int:
        n = C.c_uint64()
        self._check(self.lib.kpt_particles_count(self.h, C.byref(n)))
        return n.value

    def device_name(self) -> str:
        buf = C.create_string_buffer(64)
        self._check(self.lib.kpt_device_name(self.h, buf, 64))
        return buf.value.decode()

    def set_seed(self, seed: int):
        self._check(self.lib.kpt_set_seed(self.h, seed))

    # fields
    def field_upload(self, slot, model, x_edges, y_edges, planes):
        xe = np.ascontiguousarray(x_edges, dtype=np.float64)
        ye = np.ascontiguousarray(y_edges, dtype=np.float64)
        ps = [np.ascontiguousarray(a, dtype=np.float64) for a in planes]
        self._check(self.lib.kpt_field_upload(self.h, slot, model, len(xe) - 1, len(ye) - 1,
                                              _dp(xe), _dp(ye), *[_dp(a) for a in ps]))

    def field_load_compact(self, path, slot, model, allow_stretched=False):
        self._check(self.lib.kpt_field_load_compact_ex(self.h, path.encode(), slot, model, int(allow_stretched)))

    def field_advance(self):
        self._check(self.lib.kpt_field_advance(self.h))

    def field_stage_compact(self, path, model, allow_stretched=False):
        """Start loading a compact-field file into the staging slot in the background."""
        self._check(self.lib.kpt_field_stage_compact(self.h, path.encode(), model, int(allow_stretched)))

    def field_advance_staged(self):
        """Finish the stage and rotate lower <- upper <- staged; returns the file read seconds."""
        t = C.c_double(0.0)
        self._check(self.lib.kpt_field_advance_staged(self.h, C.byref(t)))
        return t.value

    def field_coefficients(self, slot, nx, ny, model):
        out = np.empty((ny + 2) * (nx + 2) * model)
        self._check(self.lib.kpt_field_download_coefficients(self.h, slot, _dp(out), out.size))
        return out.reshape(ny + 2, nx + 2, model)

    # particles
    def upload(self, parts: dict, count=None):
        n = len(parts["x"]) if count is None else count
        h = particles_struct(parts)
        self._check(self.lib.kpt_particles_upload(self.h, C.byref(h), n))

    def download(self, count=None, system=False) -> dict:
        """The stored particles; `system` adds the previous positions and sort keys."""
        n = self.count if count is None else count
        parts = empty_particles(n, system=system)
        h = particles_struct(parts)
        self._check(self.lib.kpt_particles_download(self.h, C.byref(h), n))
        return parts

    def inject(self, s: DeviceSettings, n, t0, frame, opts: InjectOpts | None = None) -> int:
        out = C.c_uint64()
        self._check(self.lib.kpt_inject(self.h, C.byref(s), C.byref(opts) if opts else None, n,
                                        t0, frame, C.byref(out)))
        return out.value

    def push_interval(self, s: DeviceSettings, t0, frame=0) -> PushStats:
        st = PushStats()
        self._check(self.lib.kpt_push_interval(self.h, C.byref(s), t0, frame, C.byref(st)))
        return st

    def push_interval_noalloc(self, s: DeviceSettings, t0, frame=0, stats=True) -> PushStats | None:
        """kpt_push_interval_noalloc: the interval push with no device allocation in
        the call (stats=False passes no stats record and returns None)."""
        st = PushStats() if stats else None
        self._check(self.lib.kpt_push_interval_noalloc(self.h, C.byref(s), t0, frame,
                                                       C.byref(st) if stats else None))
        return st

    def split(self, s: DeviceSettings):
        a, b = C.c_uint64(), C.c_uint64()
        self._check(self.lib.kpt_split(self.h, C.byref(s), C.byref(a), C.byref(b)))
        return a.value, b.value

    def histogram_logp(self, bins, pmin, pmax):
        out = np.empty(bins)
        self._check(self.lib.kpt_histogram_logp(self.h, bins, pmin, pmax, _dp(out)))
        return out

    def histogram_xy(self, bx, by, x_lo, x_hi, y_lo, y_hi):
        out = np.empty(bx * by)
        self._check(self.lib.kpt_histogram_xy(self.h, bx, by, x_lo, x_hi, y_lo, y_hi, _dp(out)))
        return out.reshape(by, bx)

    def snapshot_write(self, path, charge, split_ratio, p0, resumable=False):
        fn = self.lib.kpt_snapshot_write_resumable if resumable else self.lib.kpt_snapshot_write
        self._check(fn(self.h, path.encode(), charge, split_ratio, p0))

    def snapshot_read(self, path, charge, split_ratio, restart_time, dt_min):
        self._check(self.lib.kpt_snapshot_read(self.h, path.encode(), charge, split_ratio, restart_time, dt_min))

    def sort(self, s: DeviceSettings, compact=False) -> int:
        """Sort by stencil cell (inactive last); returns the active count."""
        act = C.c_uint64()
        self._check(self.lib.kpt_sort_particles(self.h, C.byref(s), int(compact), C.byref(act)))
        return act.value

    def summary(self) -> dict:
        s = Summary()
        self._check(self.lib.kpt_summary_compute(self.h, C.byref(s)))
        return s.as_dict()

    def set_noise(self, draws: np.ndarray):
        d = np.ascontiguousarray(draws, dtype=np.float64)
        assert d.ndim == 3 and d.shape[2] == 4
        self._check(self.lib.kpt_set_noise(self.h, _dp(d), d.shape[0], d.shape[1]))

    def clear_noise(self):
        self._check(self.lib.kpt_clear_noise(self.h))

    # ---- generic Parker solver ------------------------------------------------
    def generic_build(self, params: "GenericParkerParams", slot=0):
        self._check(self.lib.kpt_generic_parker_build(self.h, C.byref(params), slot))

    def generic_coefficients(self, slot, nx, ny):
        out = np.empty((ny + 2) * (nx + 2) * 18)
        self._check(self.lib.kpt_generic_parker_download(self.h, slot, _dp(out), out.size))
        return out.reshape(ny + 2, nx + 2, 18)

    def generic_time_step(self, params: "GenericParkerParams", slot=0) -> float:
        dt = C.c_double()
        self._check(self.lib.kpt_generic_parker_time_step(self.h, C.byref(params), slot, C.byref(dt)))
        return dt.value

    def generic_step(self, params: "GenericParkerParams", step, slot=0):
        """One advance_stochastic step; returns (dt, failures)."""
        dt, nf = C.c_double(), C.c_uint64()
        self._check(self.lib.kpt_generic_parker_step(self.h, C.byref(params), slot, step, C.byref(dt),
                                                     C.byref(nf)))
        return dt.value, nf.value

    # ---- generic focused-transport solver --------------------------------------
    def focus_build(self, params: "GenericFocusParams", slot=0, dvdt=None):
        d = None if dvdt is None else np.ascontiguousarray(np.stack([np.asarray(a, np.float64) for a in dvdt]))
        self._check(self.lib.kpt_generic_focus_build(self.h, C.byref(params), slot, _dp(d)))

    def focus_coefficients(self, slot, nx, ny):
        out = np.empty((ny + 2) * (nx + 2) * 27)
        self._check(self.lib.kpt_generic_focus_download(self.h, slot, _dp(out), out.size))
        return out.reshape(ny + 2, nx + 2, 27)

    def focus_time_step(self, params: "GenericFocusParams", slot=0) -> float:
        dt = C.c_double()
        self._check(self.lib.kpt_generic_focus_time_step(self.h, C.byref(params), slot, C.byref(dt)))
        return dt.value

    def focus_step(self, params: "GenericFocusParams", step, slot=0):
        """One FocusTransportSolver::advance_stochastic step; returns (dt, failures)."""
        dt, nf = C.c_double(), C.c_uint64()
        self._check(self.lib.kpt_generic_focus_step(self.h, C.byref(params), slot, step, C.byref(dt),
                                                    C.byref(nf)))
        return dt.value, nf.value

    def generic_advance_deterministic(self, params: "GenericParkerParams", slot=0):
        """ParkerSolver::advance_deterministic; returns (dt, failures)."""
        dt, nf = C.c_double(), C.c_uint64()
        self._check(self.lib.kpt_generic_parker_advance_deterministic(self.h, C.byref(params), slot, C.byref(dt),
                                                                      C.byref(nf)))
        return dt.value, nf.value

    def focus_advance_deterministic(self, params: "GenericFocusParams", slot=0):
        """FocusTransportSolver::advance_deterministic; returns (dt, failures)."""
        dt, nf = C.c_double(), C.c_uint64()
        self._check(self.lib.kpt_generic_focus_advance_deterministic(self.h, C.byref(params), slot, C.byref(dt),
                                                                     C.byref(nf)))
        return dt.value, nf.value

    # ---- ParticleSystem lifecycle ---------------------------------------------
    def capture_previous_positions(self):
        self._check(self.lib.kpt_capture_previous_positions(self.h))

    def sort_by_key(self, min_key, max_key):
        self._check(self.lib.kpt_sort_by_key(self.h, min_key, max_key))

    def compact_active_particles(self) -> int:
        act = C.c_uint64()
        self._check(self.lib.kpt_compact_active_particles(self.h, C.byref(act)))
        return act.value

    def system_snapshot_write(self, path, props: "ParticleProperties", species, policy: "EnergySplitPolicy"):
        self._check(self.lib.kpt_particle_system_write_snapshot(self.h, path.encode(), C.byref(props), species,
                                                                C.byref(policy)))

    def system_snapshot_read(self, path):
        """Replace the stored set from a ParticleSystem snapshot; returns (properties, species, policy)."""
        pr, po, sp = ParticleProperties(), EnergySplitPolicy(), C.c_int32()
        self._check(self.lib.kpt_particle_system_read_snapshot(self.h, path.encode(), C.byref(pr), C.byref(sp),
                                                               C.byref(po)))
        return pr, sp.value, po
    def split_energy_growth(self, props: "ParticleProperties", policy: "EnergySplitPolicy") -> dict:
        st = SplitStatistics()
        self._check(self.lib.kpt_split_energy_growth(self.h, C.byref(props), C.byref(policy), C.byref(st)))
        return st.as_dict()

    def reset_split_baselines(self, props: "ParticleProperties"):
        self._check(self.lib.kpt_reset_split_baselines(self.h, C.byref(props)))

    def apply_boundaries(self, slot, periodic_mask) -> int:
        out = C.c_uint64()
        self._check(self.lib.kpt_apply_boundaries(self.h, slot, periodic_mask, C.byref(out)))
        return out.value

    # ---- generic particle generator (include/ParticleGenerator.hpp) ------------
    def candidates_from_field(self, slot, pred, region=None) -> "Candidates":
        """Field-threshold candidate set; pred = (component, comparison, threshold)."""
        h = C.c_void_p()
        fp = FieldPredicate(*pred)
        self._check(self.lib.kpt_candidates_from_field(self.h, slot, C.byref(Region(*region)) if region else None,
                                                       C.byref(fp), C.byref(h)))
        return Candidates(self, h)

    def candidates_from_grid(self, slot) -> "Candidates":
        h = C.c_void_p()
        self._check(self.lib.kpt_candidates_from_grid(self.h, slot, C.byref(h)))
        return Candidates(self, h)

    def inject_from_candidates(self, cands, n, kin, props=None) -> int:
        out = C.c_uint64()
        self._check(self.lib.kpt_inject_from_candidates(self.h, cands.h, n, C.byref(kin),
                                                        C.byref(props) if props else None, C.byref(out)))
        return out.value

    def inject_uniform_if(self, slot, n, kin, region=None, pred=None, max_attempts=64, props=None):
        """(injected, rejected)"""
        a, b = C.c_uint64(), C.c_uint64()
        self._check(self.lib.kpt_inject_uniform_if(self.h, slot, C.byref(Region(*region)) if region else None,
                                                   C.byref(FieldPredicate(*pred)) if pred else None, max_attempts,
                                                   n, C.byref(kin), C.byref(props) if props else None,
                                                   C.byref(a), C.byref(b)))
        return a.value, b.value

    def recycle_uniform_if(self, slot, kin, selection, region=None, pred=None, max_attempts=64, props=None):
        """(recycled, rejected)"""
        a, b = C.c_uint64(), C.c_uint64()
        self._check(self.lib.kpt_recycle_uniform_if(self.h, slot, C.byref(Region(*region)) if region else None,
                                                    C.byref(FieldPredicate(*pred)) if pred else None, max_attempts,
                                                    selection, C.byref(kin), C.byref(props) if props else None,
                                                    C.byref(a), C.byref(b)))
        return a.value, b.value

    def recycle_from_candidates(self, cands, kin, selection, props=None) -> int:
        out = C.c_uint64()
        self._check(self.lib.kpt_recycle_from_candidates(self.h, cands.h, selection, C.byref(kin),
                                                         C.byref(props) if props else None, C.byref(out)))
        return out.value

    def generator_calls(self) -> int:
        out = C.c_uint32()
        self._check(self.lib.kpt_generator_calls(self.h, C.byref(out)))
        return out.value

    def kinematics(self, props: "ParticleProperties", count=None):
        """(gamma, speed, kinetic energy) of the first `count` stored particles."""
        n = self.count if count is None else count
        g, v, e = np.empty(n), np.empty(n), np.empty(n)
        self._check(self.lib.kpt_particle_kinematics(self.h, C.byref(props), _dp(g), _dp(v), _dp(e), n))
        return g, v, e

    # ---- learned pitch-angle closure -----------------------------------------
    def mlp_train(self, opts: "MlpTrainOpts | None" = None) -> dict:
        o = opts or mlp_default_train_opts()
        r = MlpReport()
        self._check(self.lib.kpt_mlp_train(self.h, C.byref(o), C.byref(r)))
        return r.as_dict()

    def mlp_set(self, params, hull: MlpHull):
        w = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.kpt_mlp_set(self.h, _fp(w), w.size, C.byref(hull)))

    def mlp_get(self):
        w = np.empty(MLP_PARAMS, dtype=np.float32)
        h = MlpHull()
        self._check(self.lib.kpt_mlp_get(self.h, _fp(w), w.size, C.byref(h)))
        return w, h

    def mlp_gradient(self, feats, targets):
        x = np.ascontiguousarray(feats, dtype=np.float32)
        y = np.ascontiguousarray(targets, dtype=np.float32)
        assert x.ndim == 2 and x.shape[1] == 3 and y.shape == (x.shape[0], 2)
        g = np.empty(MLP_PARAMS, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.kpt_mlp_gradient(self.h, _fp(x), _fp(y), x.shape[0], _fp(g),
                                              C.byref(loss)))
        return g, loss.value

    def mlp_infer(self, feats, mfma=True):
        x = np.ascontiguousarray(feats, dtype=np.float32)
        assert x.ndim == 2 and x.shape[1] == 3
        out = np.empty((x.shape[0], 2), dtype=np.float32)
        self._check(self.lib.kpt_mlp_infer(self.h, _fp(x), x.shape[0], 1 if mfma else 0, _fp(out)))
        return out

    def snapshot_write(self, path, charge, split_ratio, p0):
        self._check(self.lib.kpt_snapshot_write(self.h, path.encode(), charge, split_ratio, p0))

    def snapshot_read(self, path, charge, split_ratio, restart_time, dt_min):
        self._check(self.lib.kpt_snapshot_read(self.h, path.encode(), charge, split_ratio,
                                               restart_time, dt_min))


class Ensemble:
    """Batched ensemble of independent runs on one context (kpt_ensemble_*).
    Destroy it (or leave its `with` block) before the context."""

    def __init__(self, ctx: "Context", capacities):
        self.ctx, self.lib = ctx, ctx.lib
        caps = (C.c_uint64 * len(capacities))(*[int(c) for c in capacities])
        self.runs = len(capacities)
        self.h = C.c_void_p()
        ctx._check(self.lib.kpt_ensemble_create(ctx.h, self.runs, caps, C.byref(self.h)))

    def close(self):
        if self.h:
            self.lib.kpt_ensemble_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        self.ctx._check(st)

    def field_upload(self, run, model, x_edges, y_edges, planes):
        xe = np.ascontiguousarray(x_edges, dtype=np.float64)
        ye = np.ascontiguousarray(y_edges, dtype=np.float64)
        ps = [np.ascontiguousarray(a, dtype=np.float64) for a in planes]
        self._check(self.lib.kpt_ensemble_field_upload(self.h, run, model, len(xe) - 1, len(ye) - 1, _dp(xe), _dp(ye),
                                                       *[_dp(a) for a in ps]))

    def set_run(self, run, s: "DeviceSettings", seed):
        self._check(self.lib.kpt_ensemble_set_run(self.h, run, C.byref(s), seed))

    def upload(self, run, parts: dict):
        h = particles_struct(parts)
        self._check(self.lib.kpt_ensemble_particles_upload(self.h, run, C.byref(h), len(parts["x"])))

    def count(self, run) -> int:
        n = C.c_uint64()
        self._check(self.lib.kpt_ensemble_count(self.h, run, C.byref(n)))
        return n.value

    def download(self, run) -> dict:
        n = self.count(run)
        parts = empty_particles(n)
        h = particles_struct(parts)
        self._check(self.lib.kpt_ensemble_particles_download(self.h, run, C.byref(h), n))
        return parts

    def inject(self, run, n, t0, frame) -> int:
        out = C.c_uint64()
        self._check(self.lib.kpt_ensemble_inject(self.h, run, n, t0, frame, C.byref(out)))
        return out.value

    def push(self, first_frame, frames=1, graph=False):
        """Returns (per-run stats list of dicts, kernel ms)."""
        st = (EnsembleRunStats * self.runs)()
        ms = C.c_double()
        self._check(self.lib.kpt_ensemble_push(self.h, first_frame, frames, int(graph), st, C.byref(ms)))
        return [{"particle_steps": s.particle_steps, "failures": s.failures, "active_in": s.active_in}
                for s in st], ms.value

    def autotune(self, frame, window=0.25):
        chosen = EnsembleTuning()
        table = (EnsembleTuning * 16)()
        tried = C.c_int32()
        self._check(self.lib.kpt_ensemble_autotune(self.h, frame, window, C.byref(chosen), table, 16,
                                                   C.byref(tried)))
        return chosen.as_dict(), [table[k].as_dict() for k in range(tried.value)]

    def set_tuning(self, grid_blocks, groups):
        t = EnsembleTuning(grid_blocks, groups, 0.0)
        self._check(self.lib.kpt_ensemble_set_tuning(self.h, C.byref(t)))

    def histogram_logp(self, run, bins, pmin, pmax):
        out = np.empty(bins)
        self._check(self.lib.kpt_ensemble_histogram_logp(self.h, run, bins, pmin, pmax, _dp(out)))
        return out

    def sort(self) -> int:
        """Cell-sort every run's particles; returns the active count over all runs."""
        n = C.c_uint64()
        self._check(self.lib.kpt_ensemble_sort(self.h, C.byref(n)))
        return n.value
