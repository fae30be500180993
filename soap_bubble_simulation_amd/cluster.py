"""Foam cluster dynamics on the GPU: the reference's ``FoamSimulator``
(``src/physics/foam_dynamics.rs:9-244``) over a ``BubbleCluster``
(``src/physics/foam.rs:208-399``), backed by the ``sf_cluster_*`` C ABI.

Method names follow the reference: ``step``, ``compute_forces``,
``bubble_count``, ``connection_count``, ``with_default_cluster``; the cluster
operations ``add_bubble``, ``remove_bubble``, ``merge_bubbles``,
``update_connections``, ``connections`` and ``bubbles`` are on the simulator
(the reference reaches them through ``sim.cluster``).  Physics parameters are
attributes, as the reference's public fields; a change is sent to the
device on the next call.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L

_PARAM_FIELDS = ("surface_tension", "viscosity", "density", "van_der_waals_strength",
                 "repulsion_stiffness", "damping", "coalescence_threshold")


@dataclass
class Connection:
    """BubbleConnection (foam.rs:89-104)."""
    bubble_a: int
    bubble_b: int
    contact_point: np.ndarray
    contact_normal: np.ndarray
    wall_curvature_radius: float
    wall_thickness_nm: float

    def is_flat(self) -> bool:                       # foam.rs:142-144
        return abs(self.wall_curvature_radius) > 1000.0


def default_cluster_bubbles():
    """Positions and radii of BubbleCluster::create_default_cluster
    (foam.rs:402-416)."""
    pos = np.array([(0.0, 0.0, 0.0), (0.042, 0.0, 0.0), (-0.038, 0.015, 0.0), (0.0, 0.040, 0.0),
                    (0.018, -0.038, 0.015)], np.float32)
    radius = np.array([0.025, 0.02, 0.022, 0.018, 0.02], np.float32)
    return pos, radius


def _bubble_array(pos, radius, vel=None, thickness_nm=None, ids=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    radius = np.asarray(radius, np.float32).reshape(n)
    vel = np.zeros((n, 3), np.float32) if vel is None else np.asarray(vel, np.float32).reshape(n, 3)
    th = np.full(n, 500.0, np.float32) if thickness_nm is None else np.asarray(thickness_nm, np.float32).reshape(n)
    ids = np.arange(n, dtype=np.uint32) if ids is None else np.asarray(ids, np.uint32).reshape(n)
    arr = (L.ClusterBubble * max(n, 1))()
    for k in range(n):
        b = arr[k]
        b.id = int(ids[k])
        b.position[:] = [float(x) for x in pos[k]]
        b.radius = float(radius[k])
        b.velocity[:] = [float(x) for x in vel[k]]
        b.thickness_nm = float(th[k])
        b.aspect_ratio = 1.0                                           # Bubble::new (foam.rs:52-63)
        b.refractive_index = 1.33
    return arr, n


class FoamSimulator:
    """GPU FoamSimulator.  ``pos`` [n][3] (m), ``radius`` [n] (m), optional
    ``vel`` [n][3], ``thickness_nm`` [n] and ``ids`` [n] (default 0..n-1)."""

    def __init__(self, pos=None, radius=None, vel=None, thickness_nm=None, ids=None):
        p = L.ClusterParams()
        L.check(L.lib().sf_cluster_params_default(p))
        self._p = p
        for name in _PARAM_FIELDS:
            setattr(self, name, float(getattr(p, name)))
        self.gravity = np.array(p.gravity, np.float64)
        self.coalescence_enabled = bool(p.coalescence_enabled)
        if pos is None:
            pos, radius = np.zeros((0, 3), np.float32), np.zeros(0, np.float32)
        arr, n = _bubble_array(pos, radius, vel, thickness_nm, ids)
        h = C.c_void_p()
        L.check(L.lib().sf_cluster_create(C.byref(p), n, arr, C.byref(h)))
        self._h = h

    @classmethod
    def with_default_cluster(cls) -> "FoamSimulator":       # foam_dynamics.rs:57-60
        pos, radius = default_cluster_bubbles()
        return cls(pos, radius)

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            L.lib().sf_cluster_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def _sync_params(self):
        p = self._p
        for name in _PARAM_FIELDS:
            setattr(p, name, float(getattr(self, name)))
        p.gravity[:] = [float(x) for x in np.asarray(self.gravity, np.float64).reshape(3)]
        p.coalescence_enabled = 1 if self.coalescence_enabled else 0
        L.check(L.lib().sf_cluster_set_params(self._h, C.byref(p)))

    # -- FoamSimulator ------------------------------------------------------
    def compute_forces(self) -> np.ndarray:                 # foam_dynamics.rs:61-92
        self._sync_params()
        out = np.zeros((self.bubble_count(), 3), np.float32)
        if len(out):
            L.check(L.lib().sf_cluster_compute_forces(self._h, L.fptr(out)))
        return out

    def step(self, dt: float) -> None:                      # foam_dynamics.rs:166-202
        self.step_many(1, dt)

    def step_many(self, n_steps: int, dt: float) -> None:
        self._sync_params()
        L.check(L.lib().sf_cluster_step(self._h, float(dt), int(n_steps)))

    def synchronize(self) -> None:
        L.check(L.lib().sf_cluster_synchronize(self._h))

    def kernel_timing(self, enable: bool):
        """(average ms, launches) of the force kernel since the last call."""
        ms = C.c_double()
        n = C.c_int64()
        L.check(L.lib().sf_cluster_kernel_timing(self._h, 1 if enable else 0, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def bubble_count(self) -> int:                          # foam_dynamics.rs:231-233
        n = C.c_int32()
        L.check(L.lib().sf_cluster_num_bubbles(self._h, C.byref(n)))
        return n.value

    def connection_count(self) -> int:                      # foam_dynamics.rs:236-238
        n = C.c_int32()
        L.check(L.lib().sf_cluster_num_connections(self._h, C.byref(n)))
        return n.value

    # -- BubbleCluster ------------------------------------------------------
    def bubbles(self) -> dict:
        """Arrays in vector order: ids, position, radius, velocity,
        thickness_nm, aspect_ratio, refractive_index."""
        n = self.bubble_count()
        arr = (L.ClusterBubble * max(n, 1))()
        if n:
            L.check(L.lib().sf_cluster_get_bubbles(self._h, arr))
        rows = arr[:n]
        return {
            "ids": np.array([b.id for b in rows], np.uint32),
            "position": np.array([list(b.position) for b in rows], np.float32).reshape(n, 3),
            "radius": np.array([b.radius for b in rows], np.float32),
            "velocity": np.array([list(b.velocity) for b in rows], np.float32).reshape(n, 3),
            "thickness_nm": np.array([b.thickness_nm for b in rows], np.float32),
            "aspect_ratio": np.array([b.aspect_ratio for b in rows], np.float32),
            "refractive_index": np.array([b.refractive_index for b in rows], np.float32),
        }

    def set_bubbles(self, pos, radius, vel=None, thickness_nm=None, ids=None) -> None:
        arr, n = _bubble_array(pos, radius, vel, thickness_nm, ids)
        L.check(L.lib().sf_cluster_set_bubbles(self._h, n, arr))

    def connections(self) -> List[Connection]:              # foam.rs:305-307
        n = self.connection_count()
        arr = (L.ClusterConnection * max(n, 1))()
        if n:
            L.check(L.lib().sf_cluster_get_connections(self._h, arr))
        return [Connection(r.bubble_a, r.bubble_b, np.array(r.contact_point, np.float32),
                           np.array(r.contact_normal, np.float32), r.wall_curvature_radius,
                           r.wall_thickness_nm) for r in arr[:n]]

    def neighbors(self) -> List[List[int]]:
        """Bubble::neighbors of every bubble, vector order."""
        n = self.bubble_count()
        offs = (C.c_int32 * (n + 1))()
        L.check(L.lib().sf_cluster_neighbors(self._h, offs, None))
        ids = (C.c_uint32 * max(offs[n], 1))()
        L.check(L.lib().sf_cluster_neighbors(self._h, offs, ids))
        return [list(ids[offs[k]:offs[k + 1]]) for k in range(n)]

    def update_connections(self) -> None:                   # foam.rs:321-360
        L.check(L.lib().sf_cluster_update_connections(self._h))

    def add_bubble(self, position: Sequence[float], radius: float) -> int:   # foam.rs:235-244
        pos = (C.c_float * 3)(*[float(x) for x in position])
        out = C.c_uint32()
        L.check(L.lib().sf_cluster_add_bubble(self._h, pos, float(radius), C.byref(out)))
        return out.value

    def remove_bubble(self, bubble_id: int) -> bool:        # foam.rs:249-259
        out = C.c_int32()
        L.check(L.lib().sf_cluster_remove_bubble(self._h, int(bubble_id), C.byref(out)))
        return bool(out.value)

    def merge_bubbles(self, id_a: int, id_b: int) -> int:   # foam.rs:365-399
        out = C.c_uint32()
        L.check(L.lib().sf_cluster_merge(self._h, int(id_a), int(id_b), C.byref(out)))
        return out.value

    # -- coupling to batched drainage ----------------------------------------
    def foam_bubbles(self, n_theta_min: int = 32, n_theta_max: int = 512, radius_min: float = 2.5e-3,
                     radius_max: float = 25e-3) -> List[L.Bubble]:
        """Descriptions of a drainage foam with one bubble per cluster bubble
        (radius, thickness, viscosity; latitude count from the radius)."""
        self._sync_params()
        n = self.bubble_count()
        arr = (L.Bubble * max(n, 1))()
        L.check(L.lib().sf_cluster_foam_bubbles(self._h, int(n_theta_min), int(n_theta_max), float(radius_min),
                                                float(radius_max), arr))
        return list(arr[:n])

    def thickness_from_foam(self, foam) -> None:
        """Each bubble's thickness_nm := its drainage foam bubble's mean film
        thickness."""
        L.check(L.lib().sf_cluster_thickness_from_foam(self._h, foam.handle))
