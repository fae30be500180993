"""Batched drainage of a foam: one device context for many bubbles.

The reference's foam (foam.rs:32-49, ``BubbleCluster`` at foam.rs:220) keeps
one scalar film thickness per bubble.  ``FoamDrainage`` gives every bubble of
a cluster the full thickness field of ``DrainageSimulator::step``
(drainage.rs:298-392) with the surfactant transport of
drainage.wgsl:178-201, all bubbles advanced together by the HIP library
(``sf_foam_*`` of include/soapfilm.h).  Each bubble has its own grid
(n_theta x n_phi, even n_phi), radius, viscosity and time step.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L

# Tile shapes (columns per lane, rows per band) by index, as in sf_foam_autotune.
TILE_SHAPES = ((1, 16), (1, 32), (2, 16), (2, 32), (2, 64), (4, 16), (4, 32), (4, 64), (1, 4), (1, 2))


@dataclass
class FoamBubble:
    n_theta: int
    n_phi: int
    bubble_radius: float
    viscosity: float
    h0: float = 500e-9
    dt: float = 0.0            # <= 0: CFL step of the initial state


def foam_params(critical_thickness: float = 10e-9, cfl_safety: float = 1.0,
                surfactant: bool = True) -> L.Params:
    """sf_params of a foam: the CPU physics defaults (config.rs:58-95,
    drainage.rs:263-288) with Γ transport on."""
    p = L.Params()
    L.check(L.lib().sf_params_cpu_default(p))
    p.flags = L.FLAG_SURFACTANT if surfactant else 0
    p.critical_thickness = critical_thickness
    p.cfl_safety = cfl_safety
    return p


class FoamDrainage:
    """Owns one ``sf_foam`` context."""

    def __init__(self, params: L.Params, bubbles: Sequence[FoamBubble]):
        self._lib = L.lib()
        self.params = params.copy()
        self.bubbles: List[FoamBubble] = list(bubbles)
        arr = (L.Bubble * len(self.bubbles))()
        for k, b in enumerate(self.bubbles):
            arr[k] = L.Bubble(int(b.n_theta), int(b.n_phi), float(b.bubble_radius),
                              float(b.viscosity), float(b.h0), float(b.dt))
        h = C.c_void_p()
        L.check(self._lib.sf_foam_create(self.params, len(self.bubbles), arr, C.byref(h)))
        self._h = h
        offs = (C.c_int64 * (len(self.bubbles) + 1))()
        L.check(self._lib.sf_foam_layout(self._h, offs))
        self.offsets = np.array(offs[:], dtype=np.int64)

    @property
    def handle(self):
        """The ``sf_foam*`` of this foam (for the C ABI)."""
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sf_foam_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- layout -------------------------------------------------------------
    @property
    def length(self) -> int:
        return int(self.offsets[-1])

    def split(self, flat: np.ndarray) -> List[np.ndarray]:
        """Per-bubble [n_theta, n_phi] views of a concatenated field."""
        return [flat[self.offsets[k]:self.offsets[k] + b.n_theta * b.n_phi].reshape(b.n_theta, b.n_phi)
                for k, b in enumerate(self.bubbles)]

    def join(self, fields: Sequence[np.ndarray], dtype=np.float64, fill: float = 0.0) -> np.ndarray:
        out = np.full(self.length, fill, dtype=dtype)
        for k, f in enumerate(fields):
            n = self.bubbles[k].n_theta * self.bubbles[k].n_phi
            out[self.offsets[k]:self.offsets[k] + n] = np.asarray(f, dtype=dtype).reshape(-1)
        return out

    # ---- state --------------------------------------------------------------
    def set_fields(self, h: np.ndarray, gamma: Optional[np.ndarray] = None) -> None:
        """Concatenated FP64 fields (length ``self.length``); None keeps Γ."""
        h = np.ascontiguousarray(h, np.float64)
        g = None if gamma is None else np.ascontiguousarray(gamma, np.float64)
        assert h.size == self.length and (g is None or g.size == self.length)
        L.check(self._lib.sf_foam_set_fields_f64(self._h, L.dptr(h), L.dptr(g)))

    def set_fields_f32(self, h: np.ndarray, gamma: Optional[np.ndarray] = None) -> None:
        h = np.ascontiguousarray(h, np.float32)
        g = None if gamma is None else np.ascontiguousarray(gamma, np.float32)
        assert h.size == self.length and (g is None or g.size == self.length)
        L.check(self._lib.sf_foam_set_fields(self._h, L.fptr(h), L.fptr(g)))

    def fields(self):
        h = np.empty(self.length, np.float64)
        g = np.empty(self.length, np.float64)
        L.check(self._lib.sf_foam_copy_fields_f64(self._h, L.dptr(h), L.dptr(g)))
        return h, g

    # ---- stepping -----------------------------------------------------------
    def step(self, n_steps: int = 1) -> None:
        L.check(self._lib.sf_foam_step(self._h, int(n_steps)))

    def synchronize(self) -> None:
        L.check(self._lib.sf_foam_synchronize(self._h))

    def use_graph(self, enable: bool) -> None:
        L.check(self._lib.sf_foam_use_graph(self._h, 1 if enable else 0))

    def set_steps_per_launch(self, steps: int) -> None:
        """1 or 2 steps per stencil launch (two: the state is streamed once
        per two steps, bit-identical to one-step launches)."""
        L.check(self._lib.sf_foam_set_steps_per_launch(self._h, int(steps)))

    def steps_per_launch(self) -> int:
        k = C.c_int32()
        L.check(self._lib.sf_foam_get_steps_per_launch(self._h, C.byref(k)))
        return k.value

    def autotune(self, trial_steps: int = 8, return_times: bool = False):
        """Picks the fastest tile shape per size bucket; optionally also
        returns the trial kernel times [bucket][shape] in ms (-1: empty)."""
        chosen = (C.c_int32 * 4)()
        times = np.empty(4 * len(TILE_SHAPES), np.float64)
        L.check(self._lib.sf_foam_autotune(self._h, int(trial_steps), chosen, L.dptr(times)))
        if return_times:
            return list(chosen), times.reshape(4, len(TILE_SHAPES))
        return list(chosen)

    def set_tile_shapes(self, shapes: Sequence[int]) -> None:
        arr = (C.c_int32 * 4)(*[int(v) for v in shapes])
        L.check(self._lib.sf_foam_set_tile_shapes(self._h, arr))

    def bubble_dt(self) -> np.ndarray:
        dt = np.empty(len(self.bubbles), np.float64)
        L.check(self._lib.sf_foam_bubble_dt(self._h, L.dptr(dt)))
        return dt

    def stats(self):
        n = len(self.bubbles)
        m, c, lo, hi = (np.empty(n, np.float64) for _ in range(4))
        L.check(self._lib.sf_foam_stats(self._h, L.dptr(m), L.dptr(c), L.dptr(lo), L.dptr(hi)))
        return {"mass": m, "dt_cfl": c, "min": lo, "max": hi}

    def kernel_timing(self, enable: bool):
        ms = C.c_double()
        n = C.c_int64()
        L.check(self._lib.sf_foam_kernel_timing(self._h, 1 if enable else 0, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_active(self, mask: Optional[np.ndarray]) -> None:
        """Steps only bubbles with mask[b] != 0 (None: all)."""
        m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        L.check(self._lib.sf_foam_set_active(self._h, L.u8ptr(m)))

    def set_state(self, hi: np.ndarray, lo: np.ndarray, gamma: Optional[np.ndarray] = None) -> None:
        """Float-float state (host arrays of the layout length)."""
        hi = np.ascontiguousarray(hi, np.float32)
        lo = np.ascontiguousarray(lo, np.float32)
        g = None if gamma is None else np.ascontiguousarray(gamma, np.float32)
        L.check(self._lib.sf_foam_set_state(self._h, hi.ctypes.data, lo.ctypes.data,
                                             None if g is None else g.ctypes.data))

    def resolved_bubbles(self) -> List[FoamBubble]:
        arr = (L.Bubble * len(self.bubbles))()
        L.check(self._lib.sf_foam_bubbles(self._h, arr))
        return [FoamBubble(b.n_theta, b.n_phi, b.bubble_radius, b.viscosity, b.h0, b.dt) for b in arr]

    def info(self):
        cells = C.c_int64()
        items = C.c_int32()
        var = (C.c_int32 * 4)()
        L.check(self._lib.sf_foam_info(self._h, C.byref(cells), C.byref(items), var))
        return {"cells": cells.value, "items": items.value, "variants": list(var)}
