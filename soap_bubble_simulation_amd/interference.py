"""Thin-film interference backed by the gfx950 HIP kernels.

``InterferenceCalculator`` mirrors src/physics/interference.rs:127-513 and
``generate_interference_lut`` mirrors src/render/interference_lut.rs:18-41.
Every evaluation, including the single-point helpers, runs in a HIP kernel
(FP32 arithmetic); LUT addressing is formed on the host in FP64 exactly like
the reference.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from . import _lib as L

LUT_THICKNESS_SAMPLES = 256      # interference_lut.rs:9
LUT_ANGLE_SAMPLES = 64           # interference_lut.rs:10
LUT_MAX_THICKNESS_NM = 2000.0    # interference_lut.rs:13

RED_NM, GREEN_NM, BLUE_NM = 650.0, 532.0, 450.0     # interference.rs:24-31
REFRACTIVE_INDEX_AIR = 1.0                          # interference.rs:36
REFRACTIVE_INDEX_SOAP_FILM = 1.33                   # interference.rs:38

def _lib():
    return L.lib()


def _f32(x) -> np.ndarray:
    return np.ascontiguousarray(np.atleast_1d(np.asarray(x, np.float32)))


@dataclass
class RgbColor:
    """RgbColor (interference.rs:41-94): components clamped to [0, 1]."""
    red: float
    green: float
    blue: float

    @classmethod
    def new(cls, red: float, green: float, blue: float) -> "RgbColor":
        c = lambda v: min(max(float(v), 0.0), 1.0)
        return cls(c(red), c(green), c(blue))

    def to_f32_tuple(self):
        return (float(np.float32(self.red)), float(np.float32(self.green)), float(np.float32(self.blue)))

    def to_u8_tuple(self):
        r = lambda v: int(np.floor(v * 255.0 + 0.5))          # f64::round, v >= 0
        return (r(self.red), r(self.green), r(self.blue))

    def to_f32_array_with_alpha(self, alpha: float):
        return [*self.to_f32_tuple(), float(np.float32(alpha))]


@dataclass
class FresnelCoefficients:
    reflectance_s: float
    reflectance_p: float
    reflectance_average: float


class InterferenceCalculator:
    """GPU-backed mirror of InterferenceCalculator (interference.rs:127-513)."""

    def __init__(self, refractive_index_film: float = REFRACTIVE_INDEX_SOAP_FILM,
                 wavelengths_nm: Sequence[float] = (RED_NM, GREEN_NM, BLUE_NM)):
        self._n = float(refractive_index_film)
        self._w = np.asarray(wavelengths_nm, np.float64)

    @classmethod
    def new(cls, refractive_index_film: float) -> "InterferenceCalculator":
        return cls(refractive_index_film)

    @classmethod
    def default_soap_film(cls) -> "InterferenceCalculator":
        return cls(REFRACTIVE_INDEX_SOAP_FILM)

    @classmethod
    def with_custom_wavelengths(cls, n, red_nm, green_nm, blue_nm) -> "InterferenceCalculator":
        return cls(n, (red_nm, green_nm, blue_nm))

    def refractive_index_film(self) -> float:
        return self._n

    def refractive_index_medium(self) -> float:
        return REFRACTIVE_INDEX_AIR

    # --- batched device evaluations -----------------------------------------
    def terms(self, d_nm, cos_i, lambda_nm) -> np.ndarray:
        d, c, w = np.broadcast_arrays(_f32(d_nm), _f32(cos_i), _f32(lambda_nm))
        d, c, w = (np.ascontiguousarray(a, np.float32) for a in (d, c, w))
        out = np.empty((d.size, 6), np.float32)
        L.check(_lib().sf_interference_terms(self._n, d.size, L.fptr(d), L.fptr(c), L.fptr(w),
                                             L.fptr(out)))
        return out

    def reflected_intensity(self, d_nm, cos_i, lambda_nm) -> np.ndarray:
        d, c, w = np.broadcast_arrays(_f32(d_nm), _f32(cos_i), _f32(lambda_nm))
        d, c, w = (np.ascontiguousarray(a, np.float32) for a in (d, c, w))
        out = np.empty(d.size, np.float32)
        L.check(L.lib().sf_reflected_intensity(self._n, d.size, L.fptr(d), L.fptr(c), L.fptr(w),
                                               L.fptr(out)))
        return out

    def colors(self, d_nm, cos_i, gamma: bool = True) -> np.ndarray:
        d, c = np.broadcast_arrays(_f32(d_nm), _f32(cos_i))
        d, c = (np.ascontiguousarray(a, np.float32) for a in (d, c))
        out = np.empty((d.size, 3), np.float32)
        L.check(_lib().sf_interference_color(self._n, L.dptr(self._w), d.size, L.fptr(d),
                                             L.fptr(c), int(bool(gamma)), L.fptr(out)))
        return out

    # --- reference API ------------------------------------------------------
    def calculate_transmission_angle_cos(self, cos_theta_incident: float) -> float:
        return float(self.terms(0.0, cos_theta_incident, 500.0)[0, 0])

    def optical_path(self, d_nm, cos_t, lambda_nm) -> np.ndarray:
        d, c, w = np.broadcast_arrays(_f32(d_nm), _f32(cos_t), _f32(lambda_nm))
        d, c, w = (np.ascontiguousarray(a, np.float32) for a in (d, c, w))
        out = np.empty((d.size, 2), np.float32)
        L.check(L.lib().sf_optical_path(self._n, d.size, L.fptr(d), L.fptr(c), L.fptr(w),
                                        L.fptr(out)))
        return out

    def calculate_optical_path_difference(self, film_thickness_nm, cos_theta_transmitted,
                                          wavelength_nm) -> float:
        return float(self.optical_path(film_thickness_nm, cos_theta_transmitted, wavelength_nm)[0, 0])

    def calculate_phase_difference(self, optical_path_difference_nm, wavelength_nm) -> float:
        o, w = np.broadcast_arrays(_f32(optical_path_difference_nm), _f32(wavelength_nm))
        o, w = (np.ascontiguousarray(a, np.float32) for a in (o, w))
        out = np.empty(o.size, np.float32)
        L.check(L.lib().sf_phase_difference(o.size, L.fptr(o), L.fptr(w), L.fptr(out)))
        return float(out[0])

    def calculate_fresnel_reflection(self, cos_theta_incident: float) -> FresnelCoefficients:
        t = self.terms(1.0, cos_theta_incident, 500.0)[0]
        return FresnelCoefficients(float(t[3]), float(t[4]), float(t[5]))

    def calculate_reflected_intensity(self, film_thickness_nm, cos_theta_incident,
                                      wavelength_nm) -> float:
        return float(self.reflected_intensity(film_thickness_nm, cos_theta_incident, wavelength_nm)[0])

    def calculate_interference_color(self, film_thickness_nm, cos_theta_incident) -> RgbColor:
        r = self.colors(film_thickness_nm, cos_theta_incident, True)[0]
        return RgbColor.new(*r)

    def calculate_interference_color_linear(self, film_thickness_nm, cos_theta_incident) -> RgbColor:
        r = self.colors(film_thickness_nm, cos_theta_incident, False)[0]
        return RgbColor.new(*r)

    def calculate_interference_color_with_index(self, film_thickness_nm, cos_theta_incident,
                                                refractive_index) -> RgbColor:
        return InterferenceCalculator(refractive_index).calculate_interference_color(
            film_thickness_nm, cos_theta_incident)

    def generate_color_lookup_table_array(self, min_nm, max_nm, num_samples, cos_theta) -> np.ndarray:
        out = np.empty((num_samples, 3), np.float32)
        L.check(L.lib().sf_lut_1d(self._n, L.dptr(self._w), float(min_nm), float(max_nm),
                                  int(num_samples), float(cos_theta), L.fptr(out)))
        return out

    def generate_color_lookup_table(self, min_nm, max_nm, num_samples, cos_theta) -> List[RgbColor]:
        return [RgbColor.new(*r) for r in
                self.generate_color_lookup_table_array(min_nm, max_nm, num_samples, cos_theta)]

    def generate_2d_color_lookup_table_array(self, min_nm, max_nm, thickness_samples,
                                             angle_samples, gamma: bool = True) -> np.ndarray:
        out = np.empty((thickness_samples, angle_samples, 3), np.float32)
        L.check(L.lib().sf_lut_2d(self._n, L.dptr(self._w), float(min_nm), float(max_nm),
                                  int(thickness_samples), int(angle_samples), int(bool(gamma)),
                                  L.fptr(out)))
        return out

    def generate_2d_color_lookup_table(self, min_nm, max_nm, thickness_samples,
                                       angle_samples) -> List[List[RgbColor]]:
        a = self.generate_2d_color_lookup_table_array(min_nm, max_nm, thickness_samples, angle_samples)
        return [[RgbColor.new(*px) for px in row] for row in a]


def generate_interference_lut(refractive_index: float, intensity: float) -> bytes:
    """generate_interference_lut (interference_lut.rs:18-41): RGBA8 bytes,
    angle-major 64 x 256."""
    out = np.empty(LUT_ANGLE_SAMPLES * LUT_THICKNESS_SAMPLES * 4, np.uint8)
    L.check(L.lib().sf_lut_rgba8(float(refractive_index), float(intensity), L.u8ptr(out)))
    return out.tobytes()


def interference_lut_rgba8(refractive_index: float = 1.33, intensity: float = 1.0) -> np.ndarray:
    return np.frombuffer(generate_interference_lut(refractive_index, intensity),
                         np.uint8).reshape(LUT_ANGLE_SAMPLES, LUT_THICKNESS_SAMPLES, 4)


def interference_lut_rgb(refractive_index: float = 1.33, intensity: float = 1.0) -> np.ndarray:
    """Pre-quantisation LUT values [64][256][3]."""
    out = np.empty((LUT_ANGLE_SAMPLES, LUT_THICKNESS_SAMPLES, 3), np.float32)
    L.check(L.lib().sf_lut_rgb_f32(float(refractive_index), float(intensity), L.fptr(out)))
    return out


def spectral_rgb(ctx, camera, cmf, refractive_index: float = 1.33) -> np.ndarray:
    """81-wavelength RGB of the context's current thickness field:
    3 planes [3][n_theta][n_phi]."""
    cam = L.Camera()
    for k in range(3):
        cam.position[k] = float(camera[k])
    cm = np.ascontiguousarray(cmf, np.float32).reshape(81 * 3)
    out = np.empty((3, ctx.n_theta, ctx.n_phi), np.float32)
    L.check(L.lib().sf_spectral_rgb(ctx.handle, C.byref(cam), float(refractive_index),
                                    L.fptr(cm), L.fptr(out)))
    return out
