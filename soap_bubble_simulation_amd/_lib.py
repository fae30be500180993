"""ctypes binding of ``libsoapfilm.so`` (the C ABI of ``include/soapfilm.h``).

The library is built in-tree by ``soap_bubble_simulation_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback for any compute
entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libsoapfilm.so")

SF_OK = 0
SF_ERR_INVALID = -1
SF_ERR_DEVICE = -2
SF_ERR_NOMEM = -3
SF_ERR_INTERNAL = -4

FLAG_MARANGONI = 1 << 0
FLAG_SURFACTANT = 1 << 1
FLAG_TFE = 1 << 2
FLAG_POLES_COPY_H = 1 << 3
FLAG_POLES_AVG_GAMMA = 1 << 4
FLAG_DIAG = 1 << 5
SPECTRAL_OPERAND_AUTO = 0
SPECTRAL_OPERAND_BF16_PAIR = 1
SPECTRAL_OPERAND_FP16 = 2

PREDICT_NETWORK = 0
PREDICT_IDENTITY = 1
PREDICT_EXTRAPOLATE = 2
NOALLOC_MAX_THETA = 16384
FNO_PATH_FUSED_ROWS = 1 << 0
FNO_PATH_FUSED_THETA = 1 << 1
FNO_PATH_LIFT_DFT = 1 << 2
FNO_PATH_WAVE_ROWS = 1 << 3


class SoapFilmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("soapfilm error %d: %s" % (code, msg))
        self.code = code


class Params(C.Structure):
    """Mirror of ``sf_params``."""
    _fields_ = [
        ("dt", C.c_double),
        ("gravity", C.c_double),
        ("viscosity", C.c_double),
        ("density", C.c_double),
        ("diffusion_coeff", C.c_double),
        ("bubble_radius", C.c_double),
        ("critical_thickness", C.c_double),
        ("gamma_air", C.c_double),
        ("gamma_reduction", C.c_double),
        ("surfactant_diffusion", C.c_double),
        ("marangoni_coeff", C.c_double),
        ("sin_guard", C.c_double),
        ("tfe_thickness_ratio", C.c_double),
        ("tfe_first_spawn", C.c_double),
        ("tfe_spawn_interval", C.c_double),
        ("cfl_safety", C.c_double),
        ("flags", C.c_uint32),
        ("rebase_interval", C.c_uint32),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.pointer(p)[0] = self
        return p


class Stats(C.Structure):
    """Mirror of ``sf_stats``."""
    _fields_ = [
        ("min_thickness", C.c_double),
        ("max_thickness", C.c_double),
        ("mass", C.c_double),
        ("dt_cfl", C.c_double),
        ("current_time", C.c_double),
        ("n_ruptured", C.c_int64),
        ("n_tfe", C.c_int32),
        ("has_critical_region", C.c_int32),
        ("steps", C.c_int64),
    ]


class Camera(C.Structure):
    _fields_ = [("position", C.c_double * 3)]


class Bubble(C.Structure):
    """Mirror of ``sf_bubble`` (one bubble of a foam)."""
    _fields_ = [
        ("n_theta", C.c_int32),
        ("n_phi", C.c_int32),
        ("bubble_radius", C.c_double),
        ("viscosity", C.c_double),
        ("h0", C.c_double),
        ("dt", C.c_double),
    ]


class CausticParams(C.Structure):
    """Mirror of ``sf_caustic_params``."""
    _fields_ = [
        ("refractive_index", C.c_float), ("film_thickness_scale", C.c_float), ("light_dir", C.c_float * 3),
        ("light_intensity", C.c_float), ("focal_length", C.c_float), ("caustic_intensity", C.c_float),
        ("caustic_sharpness", C.c_float), ("branch_threshold", C.c_float), ("ground_y", C.c_float),
        ("bubble_radius", C.c_float), ("time", C.c_float),
    ]


class FnoConfig(C.Structure):
    """Mirror of ``sf_fno_config``."""
    _fields_ = [
        ("n_theta", C.c_int32), ("n_phi", C.c_int32), ("width", C.c_int32),
        ("modes_theta", C.c_int32), ("modes_phi", C.c_int32), ("layers", C.c_int32),
        ("window", C.c_int32), ("max_batch", C.c_int32), ("seed", C.c_uint64),
    ]


class FnoStats(C.Structure):
    """Mirror of ``sf_fno_stats``."""
    _fields_ = [
        ("windows", C.c_int64), ("bubble_windows", C.c_int64), ("accepted", C.c_int64),
        ("fallbacks", C.c_int64), ("max_residual", C.c_double), ("sum_residual", C.c_double),
        ("max_increment_residual", C.c_double), ("surrogate_ms", C.c_double),
        ("judge_ms", C.c_double), ("fallback_ms", C.c_double),
    ]


class ClusterParams(C.Structure):
    """Mirror of ``sf_cluster_params``."""
    _fields_ = [
        ("surface_tension", C.c_float), ("viscosity", C.c_float), ("density", C.c_float),
        ("gravity", C.c_float * 3), ("van_der_waals_strength", C.c_float),
        ("repulsion_stiffness", C.c_float), ("damping", C.c_float),
        ("coalescence_enabled", C.c_int32), ("coalescence_threshold", C.c_float),
    ]


class ClusterBubble(C.Structure):
    """Mirror of ``sf_cluster_bubble``."""
    _fields_ = [
        ("id", C.c_uint32), ("position", C.c_float * 3), ("radius", C.c_float),
        ("velocity", C.c_float * 3), ("thickness_nm", C.c_float), ("aspect_ratio", C.c_float),
        ("refractive_index", C.c_float),
    ]


class ClusterConnection(C.Structure):
    """Mirror of ``sf_cluster_connection``."""
    _fields_ = [
        ("bubble_a", C.c_uint32), ("bubble_b", C.c_uint32), ("contact_point", C.c_float * 3),
        ("contact_normal", C.c_float * 3), ("wall_curvature_radius", C.c_float),
        ("wall_thickness_nm", C.c_float),
    ]


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol of include/soapfilm.h.
SIGNATURES = {
    "sf_last_error": (C.c_char_p, []),
    "sf_version": (C.c_char_p, []),
    "sf_device_available": (C.c_int, []),
    "sf_params_cpu_default": (C.c_int, [C.POINTER(Params)]),
    "sf_params_gpu_default": (C.c_int, [C.POINTER(Params)]),
    "sf_drainage_create": (C.c_int, [C.POINTER(Params), C.c_int32, C.c_int32, C.c_double,
                                     C.c_double, C.POINTER(_vp)]),
    "sf_drainage_destroy": (None, [_vp]),
    "sf_drainage_set_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "sf_drainage_get_params": (C.c_int, [_vp, C.POINTER(Params)]),
    "sf_drainage_set_fields": (C.c_int, [_vp, _fp, _fp]),
    "sf_drainage_set_fields_f64": (C.c_int, [_vp, _dp, _dp]),
    "sf_drainage_step": (C.c_int, [_vp, C.c_int32, C.c_double]),
    "sf_drainage_copy_field": (C.c_int, [_vp, _fp, _fp]),
    "sf_drainage_copy_field_f64": (C.c_int, [_vp, _dp, _dp]),
    "sf_drainage_get_thickness": (C.c_int, [_vp, C.c_double, C.c_double, _dp]),
    "sf_drainage_get_thickness_batch": (C.c_int, [_vp, C.c_int32, _dp, _dp, _dp]),
    "sf_drainage_stats": (C.c_int, [_vp, C.POINTER(Stats)]),
    "sf_field_stats_noalloc": (C.c_int, [C.POINTER(Params), C.c_int32, C.c_int32, _vp, _vp, _vp, _vp]),
    "sf_drainage_reset": (C.c_int, [_vp, C.c_double, C.c_double]),
    "sf_drainage_set_marginal_regeneration": (C.c_int, [_vp, C.c_int32]),
    "sf_drainage_set_tfe_thickness_ratio": (C.c_int, [_vp, C.c_double]),
    "sf_drainage_active_tfe_count": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_drainage_current_time": (C.c_int, [_vp, _dp]),
    "sf_drainage_history": (C.c_int, [_vp, C.c_int32, _dp, _dp, C.POINTER(C.c_int32)]),
    "sf_drainage_tfe_mask": (C.c_int, [_vp, _u8p]),
    "sf_drainage_rupture_mask": (C.c_int, [_vp, _u8p]),
    "sf_drainage_device_fields": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "sf_drainage_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sf_drainage_kernel_timing": (C.c_int, [_vp, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    "sf_lut_rgba8": (C.c_int, [C.c_float, C.c_float, _u8p]),
    "sf_lut_rgb_f32": (C.c_int, [C.c_float, C.c_float, _fp]),
    "sf_lut_2d": (C.c_int, [C.c_double, _dp, C.c_double, C.c_double, C.c_int32, C.c_int32,
                            C.c_int32, _fp]),
    "sf_lut_1d": (C.c_int, [C.c_double, _dp, C.c_double, C.c_double, C.c_int32, C.c_double, _fp]),
    "sf_reflected_intensity": (C.c_int, [C.c_double, C.c_int32, _fp, _fp, _fp, _fp]),
    "sf_interference_color": (C.c_int, [C.c_double, _dp, C.c_int32, _fp, _fp, C.c_int32, _fp]),
    "sf_interference_terms": (C.c_int, [C.c_double, C.c_int32, _fp, _fp, _fp, _fp]),
    "sf_optical_path": (C.c_int, [C.c_double, C.c_int32, _fp, _fp, _fp, _fp]),
    "sf_phase_difference": (C.c_int, [C.c_int32, _fp, _fp, _fp]),
    "sf_spectral_kernel_timing": (C.c_int, [_vp, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    "sf_drainage_set_steps_per_launch": (C.c_int, [_vp, C.c_int32]),
    "sf_drainage_get_steps_per_launch": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_spectral_set_operand": (C.c_int, [_vp, C.c_int32]),
    "sf_spectral_get_operand": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_drainage_step_spectral": (C.c_int, [_vp, C.c_int32, C.c_double, C.POINTER(Camera),
                                            C.c_double, _fp]),
    "sf_drainage_step_spectral_frames": (C.c_int, [_vp, C.c_int32, C.c_double, C.POINTER(Camera),
                                                   C.c_double, _fp, _vp]),
    "sf_drainage_synchronize": (C.c_int, [_vp]),
    "sf_spectral_rgb": (C.c_int, [_vp, C.POINTER(Camera), C.c_double, _fp, _fp]),
    "sf_spectral_device_rgb": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sf_spectral_copy_rgb": (C.c_int, [_vp, _fp]),
    "sf_render_lut_rgba8": (C.c_int, [_vp, _u8p, C.c_double, C.c_float, C.POINTER(Camera), _u8p]),
    "sf_caustic_params_default": (C.c_int, [C.POINTER(CausticParams)]),
    "sf_caustics": (C.c_int, [_vp, C.POINTER(CausticParams), _fp]),
    "sf_caustics_device_map": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sf_caustics_field": (C.c_int, [C.POINTER(CausticParams), C.c_int32, C.c_int32, _fp, _fp]),
    "sf_foam_create": (C.c_int, [C.POINTER(Params), C.c_int32, C.POINTER(Bubble), C.POINTER(_vp)]),
    "sf_foam_destroy": (None, [_vp]),
    "sf_foam_layout": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "sf_foam_set_fields": (C.c_int, [_vp, _fp, _fp]),
    "sf_foam_set_fields_f64": (C.c_int, [_vp, _dp, _dp]),
    "sf_foam_copy_fields_f64": (C.c_int, [_vp, _dp, _dp]),
    "sf_foam_step": (C.c_int, [_vp, C.c_int32]),
    "sf_foam_synchronize": (C.c_int, [_vp]),
    "sf_foam_use_graph": (C.c_int, [_vp, C.c_int32]),
    "sf_foam_set_steps_per_launch": (C.c_int, [_vp, C.c_int32]),
    "sf_foam_get_steps_per_launch": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_foam_autotune": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_int32), _dp]),
    "sf_foam_set_tile_shapes": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_foam_bubble_dt": (C.c_int, [_vp, _dp]),
    "sf_foam_stats": (C.c_int, [_vp, _dp, _dp, _dp, _dp]),
    "sf_foam_kernel_timing": (C.c_int, [_vp, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    "sf_foam_info": (C.c_int, [_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "sf_foam_set_active": (C.c_int, [_vp, _u8p]),
    "sf_foam_set_state": (C.c_int, [_vp, _vp, _vp, _vp]),
    "sf_foam_device_state": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "sf_foam_bubbles": (C.c_int, [_vp, C.POINTER(Bubble)]),
    "sf_foam_stream": (C.c_int, [_vp, C.POINTER(_vp)]),
    "sf_foam_num_bubbles": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_fno_config_default": (C.c_int, [C.POINTER(FnoConfig)]),
    "sf_fno_create": (C.c_int, [C.POINTER(FnoConfig), C.POINTER(_vp)]),
    "sf_fno_destroy": (None, [_vp]),
    "sf_fno_num_params": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "sf_fno_get_params": (C.c_int, [_vp, _fp]),
    "sf_fno_set_params": (C.c_int, [_vp, _fp]),
    "sf_fno_get_scales": (C.c_int, [_vp, _dp]),
    "sf_fno_set_scales": (C.c_int, [_vp, _dp]),
    "sf_fno_forward": (C.c_int, [_vp, C.c_int32, _fp, _fp]),
    "sf_fno_gradient": (C.c_int, [_vp, C.c_int32, _fp, _fp, _dp, _fp]),
    "sf_fno_kernel_paths": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_fno_train_step": (C.c_int, [_vp, C.c_int32, _fp, _fp, C.c_double, _dp]),
    "sf_fno_train_foam": (C.c_int, [_vp, _vp, C.POINTER(Params), C.c_int32, C.c_double, _dp]),
    "sf_fno_advance_foam": (C.c_int, [_vp, _vp, C.POINTER(Params), C.c_int32, C.c_double, C.c_int32,
                                      C.POINTER(FnoStats)]),
    "sf_fno_features": (C.c_int, [_vp, _vp, C.POINTER(Params), C.c_int32, C.c_int32, _fp]),
    "sf_cluster_params_default": (C.c_int, [C.POINTER(ClusterParams)]),
    "sf_cluster_create": (C.c_int, [C.POINTER(ClusterParams), C.c_int32, C.POINTER(ClusterBubble),
                                    C.POINTER(_vp)]),
    "sf_cluster_destroy": (None, [_vp]),
    "sf_cluster_set_params": (C.c_int, [_vp, C.POINTER(ClusterParams)]),
    "sf_cluster_step": (C.c_int, [_vp, C.c_float, C.c_int32]),
    "sf_cluster_synchronize": (C.c_int, [_vp]),
    "sf_cluster_kernel_timing": (C.c_int, [_vp, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    "sf_cluster_compute_forces": (C.c_int, [_vp, _fp]),
    "sf_cluster_num_bubbles": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_cluster_get_bubbles": (C.c_int, [_vp, C.POINTER(ClusterBubble)]),
    "sf_cluster_set_bubbles": (C.c_int, [_vp, C.c_int32, C.POINTER(ClusterBubble)]),
    "sf_cluster_num_connections": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sf_cluster_get_connections": (C.c_int, [_vp, C.POINTER(ClusterConnection)]),
    "sf_cluster_neighbors": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]),
    "sf_cluster_update_connections": (C.c_int, [_vp]),
    "sf_cluster_add_bubble": (C.c_int, [_vp, _fp, C.c_float, C.POINTER(C.c_uint32)]),
    "sf_cluster_remove_bubble": (C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_int32)]),
    "sf_cluster_merge": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
    "sf_cluster_foam_bubbles": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_float, C.c_float, C.POINTER(Bubble)]),
    "sf_cluster_thickness_from_foam": (C.c_int, [_vp, _vp]),
}

_lib = None


def lib():
    """Load libsoapfilm.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libsoapfilm.so is not built (%s); run "
                "`python -m soap_bubble_simulation_amd.build`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().sf_last_error().decode()


def check(rc: int) -> None:
    if rc != SF_OK:
        raise SoapFilmError(rc, last_error())


def device_available() -> bool:
    return bool(lib().sf_device_available())


def fptr(a):
    return None if a is None else a.ctypes.data_as(_fp)


def dptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def u8ptr(a):
    return None if a is None else a.ctypes.data_as(_u8p)
