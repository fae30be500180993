"""ctypes wrapper of the FP64 CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY.  The oracle is a C restatement of the reference
drainage and interference physics (``src/physics/drainage.rs``,
``src/render/shaders/drainage.wgsl``, ``src/physics/interference.rs``,
``src/render/interference_lut.rs``).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg import this module: as the checker of
the HIP path, or as the timed CPU baseline.  The shipped package never imports
it.

Pinning status of each function (what checks it independently of the HIP
path) is listed in the module docstring of ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MARANGONI = 1 << 0
SURFACTANT = 1 << 1
TFE = 1 << 2
POLES_COPY_H = 1 << 3
POLES_AVG_GAMMA = 1 << 4


class Params(C.Structure):
    _fields_ = [
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
        ("flags", C.c_uint32),
    ]


class Front(C.Structure):
    _fields_ = [("phi_center", C.c_double), ("extent", C.c_double),
                ("width", C.c_double), ("velocity", C.c_double)]


def build() -> str:
    """Compile liboracle.so from the C sources next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        u8p = C.POINTER(C.c_uint8)
        fp = C.POINTER(C.c_float)
        sig = {
            "orc_params_cpu_default": (None, [C.POINTER(Params)]),
            "orc_create": (C.c_void_p, [C.POINTER(Params), C.c_int, C.c_int, C.c_double, C.c_double]),
            "orc_destroy": (None, [C.c_void_p]),
            "orc_set_params": (None, [C.c_void_p, C.POINTER(Params)]),
            "orc_set_fields": (None, [C.c_void_p, dp, dp]),
            "orc_get_fields": (None, [C.c_void_p, dp, dp]),
            "orc_step": (None, [C.c_void_p, C.c_double]),
            "orc_get_thickness": (C.c_double, [C.c_void_p, C.c_double, C.c_double]),
            "orc_min_thickness": (C.c_double, [C.c_void_p]),
            "orc_max_thickness": (C.c_double, [C.c_void_p]),
            "orc_has_critical_region": (C.c_int, [C.c_void_p]),
            "orc_current_time": (C.c_double, [C.c_void_p]),
            "orc_reset": (None, [C.c_void_p, C.c_double, C.c_double]),
            "orc_set_marginal_regeneration": (None, [C.c_void_p, C.c_int]),
            "orc_set_tfe_thickness_ratio": (None, [C.c_void_p, C.c_double]),
            "orc_active_tfe_count": (C.c_int, [C.c_void_p]),
            "orc_get_fronts": (C.c_int, [C.c_void_p, C.POINTER(Front), C.c_int]),
            "orc_next_spawn_time": (C.c_double, [C.c_void_p]),
            "orc_tfe_mask": (None, [C.c_void_p, u8p]),
            "orc_rupture_mask": (None, [C.c_void_p, u8p]),
            "orc_mass": (C.c_double, [C.c_void_p]),
            "orc_cfl_dt": (C.c_double, [C.c_void_p]),
            "orc_cell_area": (C.c_double, [C.c_int, C.c_int, C.c_double, C.c_int]),
            "orc_transmission_cos": (C.c_double, [C.c_double, C.c_double, C.c_double]),
            "orc_optical_path_difference": (C.c_double, [C.c_double] * 4),
            "orc_phase_difference": (C.c_double, [C.c_double] * 2),
            "orc_fresnel": (None, [C.c_double, C.c_double, C.c_double, dp, dp, dp]),
            "orc_reflected_intensity": (C.c_double, [C.c_double] * 5),
            "orc_interference_color": (None, [C.c_double, dp, C.c_double, C.c_double, C.c_int, dp]),
            "orc_lut_1d": (None, [C.c_double, dp, C.c_double, C.c_double, C.c_int, C.c_double, dp]),
            "orc_lut_2d": (None, [C.c_double, dp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, dp]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_lut_airy_terms": (None, [C.c_float, C.c_float, C.c_float, fp, fp, fp, fp]),
            "orc_lut_thin_film": (None, [C.c_float, C.c_float, C.c_float, C.c_float, fp]),
            "orc_lut_rgb_f32": (None, [C.c_float, C.c_float, fp]),
            "orc_lut_rgba8": (None, [C.c_float, C.c_float, u8p]),
            "orc_view_cos": (C.c_double, [C.c_double, C.c_double, C.c_double, dp]),
            "orc_spectral_rgb": (None, [dp, C.c_int, C.c_int, C.c_double, dp, C.c_double, dp, dp]),
            "orc_spectral_cells": (None, [dp, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_int,
                                          C.c_double, dp, C.c_double, dp, dp]),
            "orc_render_lut": (None, [dp, C.c_int, C.c_int, C.c_double, dp, C.POINTER(C.c_uint8), C.c_double,
                                      C.POINTER(C.c_uint8)]),
            "orc_caustics": (None, [C.POINTER(C.c_float), C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                    C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_uint8)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    """Host threads of the oracle's interior rows and sampled spectral cells
    (1 by default: the reference's loop is sequential)."""
    lib().orc_set_threads(int(n))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def cpu_default_params() -> Params:
    p = Params()
    lib().orc_params_cpu_default(C.byref(p))
    return p


class Sim:
    """FP64 drainage oracle state (restates ``DrainageSimulator``)."""

    def __init__(self, params: Params, n_theta: int, n_phi: int, h0: float = 500e-9,
                 gamma0: float = 0.5):
        self.n_theta, self.n_phi = n_theta, n_phi
        self._p = Params()
        C.pointer(self._p)[0] = params
        self._h = lib().orc_create(C.byref(self._p), n_theta, n_phi, h0, gamma0)
        if not self._h:
            raise ValueError("oracle rejected grid %dx%d" % (n_theta, n_phi))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def set_params(self, p: Params):
        self._p = p
        lib().orc_set_params(self._h, C.byref(p))

    def set_fields(self, h=None, gamma=None):
        n = self.n_theta * self.n_phi
        hh = None if h is None else np.ascontiguousarray(h, dtype=np.float64).reshape(n)
        gg = None if gamma is None else np.ascontiguousarray(gamma, dtype=np.float64).reshape(n)
        lib().orc_set_fields(self._h, None if hh is None else _dp(hh), None if gg is None else _dp(gg))

    def fields(self):
        h = np.empty((self.n_theta, self.n_phi), np.float64)
        g = np.empty((self.n_theta, self.n_phi), np.float64)
        lib().orc_get_fields(self._h, _dp(h), _dp(g))
        return h, g

    def step(self, dt: float, n: int = 1):
        for _ in range(n):
            lib().orc_step(self._h, dt)

    def get_thickness(self, theta: float, phi: float) -> float:
        return lib().orc_get_thickness(self._h, theta, phi)

    def min_thickness(self):
        return lib().orc_min_thickness(self._h)

    def max_thickness(self):
        return lib().orc_max_thickness(self._h)

    def has_critical_region(self):
        return bool(lib().orc_has_critical_region(self._h))

    def current_time(self):
        return lib().orc_current_time(self._h)

    def reset(self, h0: float, gamma0: float = 0.5):
        lib().orc_reset(self._h, h0, gamma0)

    def set_marginal_regeneration(self, enabled: bool):
        lib().orc_set_marginal_regeneration(self._h, int(bool(enabled)))

    def set_tfe_thickness_ratio(self, r: float):
        lib().orc_set_tfe_thickness_ratio(self._h, r)

    def active_tfe_count(self):
        return lib().orc_active_tfe_count(self._h)

    def fronts(self):
        buf = (Front * 16)()
        n = lib().orc_get_fronts(self._h, buf, 16)
        return [(f.phi_center, f.extent, f.width, f.velocity) for f in buf[:n]]

    def next_spawn_time(self):
        return lib().orc_next_spawn_time(self._h)

    def tfe_mask(self):
        m = np.zeros((self.n_theta, self.n_phi), np.uint8)
        lib().orc_tfe_mask(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8)))
        return m

    def rupture_mask(self):
        m = np.zeros((self.n_theta, self.n_phi), np.uint8)
        lib().orc_rupture_mask(self._h, m.ctypes.data_as(C.POINTER(C.c_uint8)))
        return m

    def mass(self):
        return lib().orc_mass(self._h)

    def cfl_dt(self):
        return lib().orc_cfl_dt(self._h)


def cell_area(n_theta, n_phi, radius, i):
    return lib().orc_cell_area(n_theta, n_phi, radius, i)


# ---- interference -----------------------------------------------------------

RGB_NM = (650.0, 532.0, 450.0)


def transmission_cos(n_film, cos_i, n_medium=1.0):
    return lib().orc_transmission_cos(n_film, n_medium, cos_i)


def optical_path_difference(n_film, d_nm, cos_t, lam):
    return lib().orc_optical_path_difference(n_film, d_nm, cos_t, lam)


def phase_difference(opd, lam):
    return lib().orc_phase_difference(opd, lam)


def fresnel(n_film, cos_i, n_medium=1.0):
    rs, rp, ra = C.c_double(), C.c_double(), C.c_double()
    lib().orc_fresnel(n_film, n_medium, cos_i, C.byref(rs), C.byref(rp), C.byref(ra))
    return rs.value, rp.value, ra.value


def reflected_intensity(n_film, d_nm, cos_i, lam, n_medium=1.0):
    return lib().orc_reflected_intensity(n_film, n_medium, d_nm, cos_i, lam)


def interference_color(n_film, d_nm, cos_i, gamma=True, lambdas=RGB_NM):
    lam = np.asarray(lambdas, np.float64)
    out = np.zeros(3, np.float64)
    lib().orc_interference_color(n_film, _dp(lam), d_nm, cos_i, int(gamma), _dp(out))
    return out


def lut_1d(n_film, tmin, tmax, n, cos_theta, lambdas=RGB_NM):
    lam = np.asarray(lambdas, np.float64)
    out = np.zeros((n, 3), np.float64)
    lib().orc_lut_1d(n_film, _dp(lam), tmin, tmax, n, cos_theta, _dp(out))
    return out


def lut_2d(n_film, tmin, tmax, nt, na, gamma=True, lambdas=RGB_NM):
    lam = np.asarray(lambdas, np.float64)
    out = np.zeros((nt, na, 3), np.float64)
    lib().orc_lut_2d(n_film, _dp(lam), tmin, tmax, nt, na, int(gamma), _dp(out))
    return out


def lut_airy_terms(t_nm, cos_theta, n_film):
    """Reflectance stage of the runtime LUT at one (thickness, angle):
    (cos_t, R, F, a[7]) in FP32, a[k] the Airy reflectance at 400..700 nm."""
    ct, r, f = C.c_float(), C.c_float(), C.c_float()
    a = (C.c_float * 7)()
    lib().orc_lut_airy_terms(t_nm, cos_theta, n_film, C.byref(ct), C.byref(r), C.byref(f), a)
    return np.float32(ct.value), np.float32(r.value), np.float32(f.value), np.array(a[:], np.float32)


def lut_thin_film(t_nm, cos_theta, n_film, intensity=1.0):
    out = (C.c_float * 3)()
    lib().orc_lut_thin_film(t_nm, cos_theta, n_film, intensity, out)
    return np.array(out[:], np.float32)


def lut_rgb_f32(n_film=1.33, intensity=1.0):
    out = np.zeros((64, 256, 3), np.float32)
    lib().orc_lut_rgb_f32(n_film, intensity, out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def lut_rgba8(n_film=1.33, intensity=1.0):
    out = np.zeros((64, 256, 4), np.uint8)
    lib().orc_lut_rgba8(n_film, intensity, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out


def view_cos(theta, phi, radius, cam):
    c = np.asarray(cam, np.float64)
    return lib().orc_view_cos(theta, phi, radius, _dp(c))


def spectral_rgb(h, radius, cam, n_film, cmf):
    h = np.ascontiguousarray(h, np.float64)
    nt, npf = h.shape
    c = np.asarray(cam, np.float64)
    cm = np.ascontiguousarray(cmf, np.float64).reshape(81 * 3)
    out = np.zeros((3, nt, npf), np.float64)
    lib().orc_spectral_rgb(_dp(h), nt, npf, radius, _dp(c), n_film, _dp(cm), _dp(out))
    return out


def spectral_cells(h, cells, radius, cam, n_film, cmf):
    h = np.ascontiguousarray(h, np.float64)
    nt, npf = h.shape
    cl = np.ascontiguousarray(cells, np.int64)
    c = np.asarray(cam, np.float64)
    cm = np.ascontiguousarray(cmf, np.float64).reshape(81 * 3)
    out = np.zeros((cl.size, 3), np.float64)
    lib().orc_spectral_cells(_dp(h), nt, npf, cl.ctypes.data_as(C.POINTER(C.c_int64)), cl.size,
                             radius, _dp(c), n_film, _dp(cm), _dp(out))
    return out


def caustics(h, focal_length=0.1, caustic_intensity=2.0, caustic_sharpness=1.5, branch_threshold=0.001):
    """Caustic intensity map of an FP32 thickness grid (caustics_compute.wgsl)
    and its branch mask; defaults of CausticParams (caustics.rs:56-78)."""
    h = np.ascontiguousarray(h, np.float32)
    nt, npf = h.shape
    out = np.zeros((nt, npf), np.float32)
    br = np.zeros((nt, npf), np.uint8)
    fp = C.POINTER(C.c_float)
    lib().orc_caustics(h.ctypes.data_as(fp), nt, npf, focal_length, caustic_intensity, caustic_sharpness,
                       branch_threshold, out.ctypes.data_as(fp), br.ctypes.data_as(C.POINTER(C.c_uint8)))
    return out, br


def render_lut(h, radius, cam, lut_rgba8, intensity=4.0):
    """RGBA8 frame [n_theta][n_phi][4] of a thickness grid (metres) sampled
    through an RGBA8 LUT (bubble.wgsl:425-439)."""
    h = np.ascontiguousarray(h, np.float64)
    nt, npf = h.shape
    lut = np.ascontiguousarray(lut_rgba8, np.uint8)
    out = np.zeros((nt, npf, 4), np.uint8)
    c = np.asarray(cam, np.float64)
    u8 = C.POINTER(C.c_uint8)
    lib().orc_render_lut(_dp(h), nt, npf, radius, _dp(c), lut.ctypes.data_as(u8), intensity,
                         out.ctypes.data_as(u8))
    return out
