"""Caustic-intensity map of the film thickness (CausticRenderer, caustics.rs).

``caustic_map`` evaluates caustics_compute.wgsl:27-126 on a thickness grid
through the HIP kernel of ``sf_caustics_field``; ``DrainageContext.caustics``
does it on a context's current state without leaving the device.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import _lib as L


def default_params() -> L.CausticParams:
    """CausticParams::default (caustics.rs:56-78)."""
    p = L.CausticParams()
    L.check(L.lib().sf_caustic_params_default(p))
    return p


def caustic_map(thickness: np.ndarray, params: Optional[L.CausticParams] = None) -> np.ndarray:
    h = np.ascontiguousarray(thickness, np.float32)
    nt, npf = h.shape
    p = params if params is not None else default_params()
    out = np.empty((nt, npf), np.float32)
    L.check(L.lib().sf_caustics_field(p, nt, npf, L.fptr(h), L.fptr(out)))
    return out
