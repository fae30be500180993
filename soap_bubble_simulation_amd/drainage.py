"""Drainage simulators backed by the gfx950 HIP kernels.

``DrainageSimulator`` mirrors the reference's CPU solver API
(src/physics/drainage.rs:212-614) and ``GPUDrainageSimulator`` its wgpu
compute-shader host (src/render/gpu_drainage.rs:81-491).  Both drive the same
C ABI context (``include/soapfilm.h``); every step, query and reduction runs
on the GPU.  Constructing either raises ``SoapFilmError`` when no GPU is
present.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import _lib as L
from .config import SimulationConfig


class DrainageContext:
    """Owner of one ``sf_drainage`` handle (one bubble, one HIP stream)."""

    def __init__(self, params: L.Params, n_theta: int, n_phi: int, h0: float,
                 gamma0: float = 0.5):
        self._lib = L.lib()
        self.n_theta, self.n_phi = int(n_theta), int(n_phi)
        h = C.c_void_p()
        L.check(self._lib.sf_drainage_create(C.byref(params), self.n_theta, self.n_phi,
                                             float(h0), float(gamma0), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sf_drainage_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # --- parameters ---------------------------------------------------------
    def params(self) -> L.Params:
        p = L.Params()
        L.check(self._lib.sf_drainage_get_params(self._h, C.byref(p)))
        return p

    def set_params(self, p: L.Params) -> None:
        L.check(self._lib.sf_drainage_set_params(self._h, C.byref(p)))

    # --- fields -------------------------------------------------------------
    def set_fields(self, h=None, gamma=None) -> None:
        n = self.n_theta * self.n_phi
        hh = None if h is None else np.ascontiguousarray(h, np.float64).reshape(n)
        gg = None if gamma is None else np.ascontiguousarray(gamma, np.float64).reshape(n)
        L.check(self._lib.sf_drainage_set_fields_f64(self._h, L.dptr(hh), L.dptr(gg)))

    def set_fields_f32(self, h=None, gamma=None) -> None:
        n = self.n_theta * self.n_phi
        hh = None if h is None else np.ascontiguousarray(h, np.float32).reshape(n)
        gg = None if gamma is None else np.ascontiguousarray(gamma, np.float32).reshape(n)
        L.check(self._lib.sf_drainage_set_fields(self._h, L.fptr(hh), L.fptr(gg)))

    def fields(self) -> Tuple[np.ndarray, np.ndarray]:
        h = np.empty((self.n_theta, self.n_phi), np.float64)
        g = np.empty((self.n_theta, self.n_phi), np.float64)
        L.check(self._lib.sf_drainage_copy_field_f64(self._h, L.dptr(h), L.dptr(g)))
        return h, g

    def fields_f32(self) -> Tuple[np.ndarray, np.ndarray]:
        h = np.empty((self.n_theta, self.n_phi), np.float32)
        g = np.empty((self.n_theta, self.n_phi), np.float32)
        L.check(self._lib.sf_drainage_copy_field(self._h, L.fptr(h), L.fptr(g)))
        return h, g

    # --- stepping -----------------------------------------------------------
    def step(self, n_steps: int = 1, dt: float = 0.0) -> None:
        L.check(self._lib.sf_drainage_step(self._h, int(n_steps), float(dt)))

    def get_thickness(self, theta: float, phi: float) -> float:
        out = C.c_double()
        L.check(self._lib.sf_drainage_get_thickness(self._h, float(theta), float(phi),
                                                     C.byref(out)))
        return out.value

    def get_thickness_batch(self, theta, phi) -> np.ndarray:
        t = np.ascontiguousarray(theta, np.float64).ravel()
        p = np.ascontiguousarray(phi, np.float64).ravel()
        if t.shape != p.shape:
            raise ValueError("theta and phi must have the same length")
        out = np.empty_like(t)
        L.check(self._lib.sf_drainage_get_thickness_batch(self._h, t.size, L.dptr(t), L.dptr(p),
                                                           L.dptr(out)))
        return out

    def stats(self) -> L.Stats:
        s = L.Stats()
        L.check(self._lib.sf_drainage_stats(self._h, C.byref(s)))
        return s

    def reset(self, h0: float, gamma0: float = 0.5) -> None:
        L.check(self._lib.sf_drainage_reset(self._h, float(h0), float(gamma0)))

    def set_marginal_regeneration(self, enabled: bool) -> None:
        L.check(self._lib.sf_drainage_set_marginal_regeneration(self._h, int(bool(enabled))))

    def set_tfe_thickness_ratio(self, ratio: float) -> None:
        L.check(self._lib.sf_drainage_set_tfe_thickness_ratio(self._h, float(ratio)))

    def active_tfe_count(self) -> int:
        v = C.c_int32()
        L.check(self._lib.sf_drainage_active_tfe_count(self._h, C.byref(v)))
        return v.value

    def current_time(self) -> float:
        v = C.c_double()
        L.check(self._lib.sf_drainage_current_time(self._h, C.byref(v)))
        return v.value

    def history(self, n: int = 1 << 16) -> Tuple[np.ndarray, np.ndarray]:
        m = np.empty(n, np.float64)
        c = np.empty(n, np.float64)
        k = C.c_int32()
        L.check(self._lib.sf_drainage_history(self._h, n, L.dptr(m), L.dptr(c), C.byref(k)))
        return m[:k.value].copy(), c[:k.value].copy()

    def tfe_mask(self) -> np.ndarray:
        m = np.empty((self.n_theta, self.n_phi), np.uint8)
        L.check(self._lib.sf_drainage_tfe_mask(self._h, L.u8ptr(m)))
        return m

    def rupture_mask(self) -> np.ndarray:
        m = np.empty((self.n_theta, self.n_phi), np.uint8)
        L.check(self._lib.sf_drainage_rupture_mask(self._h, L.u8ptr(m)))
        return m

    def stream(self) -> int:
        s = C.c_void_p()
        L.check(self._lib.sf_drainage_stream(self._h, C.byref(s)))
        return s.value or 0

    def caustics(self, params=None) -> np.ndarray:
        """Caustic-intensity map of the current thickness (caustics_compute.wgsl)."""
        from .caustics import default_params
        p = params if params is not None else default_params()
        out = np.empty((self.n_theta, self.n_phi), np.float32)
        L.check(self._lib.sf_caustics(self.handle, p, L.fptr(out)))
        return out

    def kernel_timing(self, enable: bool) -> Tuple[float, int]:
        """Average drainage-kernel duration (ms) over the launches since the
        previous call, measured with HIP events on the context's stream."""
        ms = C.c_double()
        n = C.c_int64()
        L.check(self._lib.sf_drainage_kernel_timing(self._h, int(bool(enable)), C.byref(ms),
                                                     C.byref(n)))
        return ms.value, n.value

    def set_steps_per_launch(self, steps: int) -> None:
        """1, or 2 for two steps per stencil launch where the flags and grid
        allow it (bit-identical results; see ``steps_per_launch``)."""
        L.check(self._lib.sf_drainage_set_steps_per_launch(self._h, int(steps)))

    def steps_per_launch(self) -> int:
        """Steps per stencil launch that ``step`` uses for this context."""
        k = C.c_int32()
        L.check(self._lib.sf_drainage_get_steps_per_launch(self._h, C.byref(k)))
        return k.value


class ThicknessField:
    """Snapshot of the device field with the accessors of ThicknessField
    (drainage.rs:96-204)."""

    def __init__(self, data: np.ndarray):
        self._data = data
        self._nt, self._np = data.shape
        self._dtheta = math.pi / (self._nt - 1)
        self._dphi = 2.0 * math.pi / self._np

    def get(self, theta_index: int, phi_index: int) -> float:
        return float(self._data[theta_index, phi_index])

    def theta_at(self, theta_index: int) -> float:
        return theta_index * self._dtheta

    def phi_at(self, phi_index: int) -> float:
        return phi_index * self._dphi

    def num_theta(self) -> int:
        return self._nt

    def num_phi(self) -> int:
        return self._np

    def delta_theta(self) -> float:
        return self._dtheta

    def delta_phi(self) -> float:
        return self._dphi

    def min_thickness(self) -> float:
        return float(np.nanmin(self._data))

    def max_thickness(self) -> float:
        return float(np.nanmax(self._data))

    def data(self) -> np.ndarray:
        return self._data.reshape(-1)


class DrainageSimulator:
    """GPU-backed mirror of ``DrainageSimulator`` (drainage.rs:212-614).

    ``new`` takes a ``SimulationConfig`` exactly as drainage.rs:263-288 does:
    n_theta = resolution, n_phi = 2 * resolution, uniform initial thickness,
    TFE fronts enabled.
    """

    def __init__(self, config: Optional[SimulationConfig] = None, *, params: Optional[L.Params] = None,
                 n_theta: Optional[int] = None, n_phi: Optional[int] = None,
                 initial_thickness: Optional[float] = None):
        config = config or SimulationConfig()
        p = params if params is not None else config.to_params()
        nt = n_theta if n_theta is not None else config.resolution
        npf = n_phi if n_phi is not None else config.resolution * 2
        h0 = initial_thickness if initial_thickness is not None else config.film_thickness_meters()
        self._gamma0 = config.fluid.surfactant_concentration
        self.ctx = DrainageContext(p, nt, npf, h0, self._gamma0)

    new = classmethod(lambda cls, config: cls(config))

    def step(self, dt: float) -> None:                          # drainage.rs:298
        if not dt > 0.0:
            raise ValueError("dt must be positive")
        self.ctx.step(1, dt)

    def step_many(self, n_steps: int, dt: float) -> None:
        self.ctx.step(n_steps, dt)

    def get_thickness(self, theta: float, phi: float) -> float:  # drainage.rs:516
        return self.ctx.get_thickness(theta, phi)

    def thickness_field(self) -> ThicknessField:                 # drainage.rs:555
        return ThicknessField(self.ctx.fields()[0])

    def current_time(self) -> float:                            # drainage.rs:560
        return self.ctx.current_time()

    def has_critical_region(self) -> bool:                      # drainage.rs:565
        return bool(self.ctx.stats().has_critical_region)

    def critical_thickness(self) -> float:                      # drainage.rs:570
        return self.ctx.params().critical_thickness

    def reset(self, initial_thickness: float) -> None:          # drainage.rs:575
        self.ctx.reset(initial_thickness, self._gamma0)

    def set_marginal_regeneration(self, enabled: bool) -> None:  # drainage.rs:588
        self.ctx.set_marginal_regeneration(enabled)

    def marginal_regeneration_enabled(self) -> bool:             # drainage.rs:596
        return bool(self.ctx.params().flags & L.FLAG_TFE)

    def set_tfe_thickness_ratio(self, ratio: float) -> None:     # drainage.rs:601
        self.ctx.set_tfe_thickness_ratio(ratio)

    def tfe_thickness_ratio(self) -> float:                     # drainage.rs:606
        return self.ctx.params().tfe_thickness_ratio

    def active_tfe_count(self) -> int:                           # drainage.rs:611
        return self.ctx.active_tfe_count()


@dataclass
class DrainageParams:
    """Mirror of the 64-byte uniform ``DrainageParams`` (gpu_drainage.rs:12-68)."""
    dt: float = 0.001
    gravity: float = 9.81
    viscosity: float = 0.001
    density: float = 1000.0
    diffusion_coeff: float = 1e-9
    bubble_radius: float = 0.025
    critical_thickness: float = 30e-9
    grid_width: int = 128
    grid_height: int = 64
    marangoni_enabled: int = 0
    gamma_air: float = 0.072
    gamma_reduction: float = 0.045
    surfactant_diffusion: float = 1e-9
    marangoni_coeff: float = 0.01

    def to_sf(self) -> L.Params:
        p = L.Params()
        L.check(L.lib().sf_params_gpu_default(p))
        f32 = lambda x: float(np.float32(x))        # the uniform holds f32 values
        p.dt = f32(self.dt)
        p.gravity = f32(self.gravity)
        p.viscosity = f32(self.viscosity)
        p.density = f32(self.density)
        p.diffusion_coeff = f32(self.diffusion_coeff)
        p.bubble_radius = f32(self.bubble_radius)
        p.critical_thickness = f32(self.critical_thickness)
        p.gamma_air = f32(self.gamma_air)
        p.gamma_reduction = f32(self.gamma_reduction)
        p.surfactant_diffusion = f32(self.surfactant_diffusion)
        p.marangoni_coeff = f32(self.marangoni_coeff)
        if self.marangoni_enabled:
            p.flags |= L.FLAG_MARANGONI
        return p


class GPUDrainageSimulator:
    """Mirror of ``GPUDrainageSimulator`` (gpu_drainage.rs:81-491).

    The kernel always steps with ``params.dt`` while the frame ``dt`` only
    advances ``current_time`` (gpu_drainage.rs:338-370).
    """

    def __init__(self, initial_thickness: float, grid_width: int, grid_height: int):
        self._params = DrainageParams(grid_width=grid_width, grid_height=grid_height)
        self.ctx = DrainageContext(self._params.to_sf(), grid_height, grid_width,
                                   float(np.float32(initial_thickness)), 0.5)
        self._current_time = 0.0
        self.enabled = False
        self.time_scale = 100.0
        self.steps_per_frame = 10
        self.marangoni_enabled = False

    def step(self, dt: float) -> None:                            # gpu_drainage.rs:338
        if not self.enabled:
            return
        scaled_dt = np.float32(dt) * np.float32(self.time_scale)
        step_dt = scaled_dt / np.float32(self.steps_per_frame)
        self.ctx.step(self.steps_per_frame, float(np.float32(self._params.dt)))
        for _ in range(self.steps_per_frame):
            self._current_time += float(step_dt)

    def update_params(self, params: DrainageParams) -> None:     # gpu_drainage.rs:377
        self._params = params
        self.ctx.set_params(params.to_sf())

    def set_dt(self, dt: float) -> None:                          # gpu_drainage.rs:383
        self._params.dt = dt
        self.ctx.set_params(self._params.to_sf())

    def set_physics(self, viscosity, density, gravity, diffusion) -> None:   # 389-402
        self._params.viscosity, self._params.density = viscosity, density
        self._params.gravity, self._params.diffusion_coeff = gravity, diffusion
        self.ctx.set_params(self._params.to_sf())

    def set_bubble_radius(self, radius: float) -> None:          # gpu_drainage.rs:405
        self._params.bubble_radius = radius
        self.ctx.set_params(self._params.to_sf())

    def reset(self, initial_thickness: float) -> None:           # gpu_drainage.rs:415
        self.ctx.reset(float(np.float32(initial_thickness)), 0.5)
        self._current_time = 0.0

    def set_marangoni_enabled(self, enabled: bool) -> None:      # gpu_drainage.rs:435
        self.marangoni_enabled = bool(enabled)
        self._params.marangoni_enabled = 1 if enabled else 0
        self.ctx.set_params(self._params.to_sf())

    def set_marangoni_params(self, gamma_air, gamma_reduction, surfactant_diffusion,
                             marangoni_coeff) -> None:            # gpu_drainage.rs:443
        self._params.gamma_air, self._params.gamma_reduction = gamma_air, gamma_reduction
        self._params.surfactant_diffusion = surfactant_diffusion
        self._params.marangoni_coeff = marangoni_coeff
        self.ctx.set_params(self._params.to_sf())

    def current_time(self) -> float:                              # gpu_drainage.rs:464
        return self._current_time

    def current_thickness(self) -> np.ndarray:                    # gpu_drainage.rs:471-480
        return self.ctx.fields_f32()[0]

    def concentration(self) -> np.ndarray:                        # gpu_drainage.rs:459
        return self.ctx.fields_f32()[1]

    def grid_size(self) -> Tuple[int, int]:                       # gpu_drainage.rs:483
        return self._params.grid_width, self._params.grid_height

    def params(self) -> DrainageParams:                           # gpu_drainage.rs:488
        return self._params
