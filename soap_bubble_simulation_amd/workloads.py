"""Seeded synthetic inputs of the benchmark and test workloads.

Shared by tests and ``bench.py`` so that the HIP path and the oracle see the
same numbers.  Nothing here is compute on the hot path.

* single bubble, 128 x 256 (the reference's default grid, config.rs:314):
  h = 500 nm (1 + 0.01 N(0,1)) per cell, Γ = 0.5.
* single bubble, 2048 x 4096: base thickness drawn from U[200, 1000] nm,
  smooth relative noise of spherical-harmonic degree <= 16 with peak
  amplitude 10 %, Γ = 0.5.
* spectral colour-matching curves: three Gaussians over the 81 wavelengths
  380-780 nm with seeded centre, width and height.
"""
from __future__ import annotations

import math
from typing import Tuple

import numpy as np

WAVELENGTHS_NM = 380.0 + 5.0 * np.arange(81)


def grid_angles(n_theta: int, n_phi: int) -> Tuple[np.ndarray, np.ndarray]:
    """theta_i = i pi/(n_theta-1), phi_j = j 2pi/n_phi (drainage.rs:121-157)."""
    dth = math.pi / (n_theta - 1)
    dph = 2.0 * math.pi / n_phi
    return np.arange(n_theta) * dth, np.arange(n_phi) * dph


def ff_exact(x: np.ndarray) -> np.ndarray:
    """Round FP64 values to the nearest float-float pair (hi + lo, both FP32)
    and return that pair's exact FP64 value, so that the GPU and the FP64
    oracle start from bit-identical thickness fields."""
    x = np.asarray(x, np.float64)
    hi = x.astype(np.float32)
    lo = (x - hi.astype(np.float64)).astype(np.float32)
    return hi.astype(np.float64) + lo.astype(np.float64)


def noise_field(n_theta: int, n_phi: int, seed: int, rel: float = 0.01,
                base: float = 500e-9) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return ff_exact(base * (1.0 + rel * rng.standard_normal((n_theta, n_phi))))


def _legendre_normalized(lmax: int, x: np.ndarray) -> np.ndarray:
    """P[l, m, k]: fully normalised associated Legendre functions at x[k]."""
    P = np.zeros((lmax + 1, lmax + 1, x.size))
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    pmm = np.full(x.size, math.sqrt(1.0 / (4.0 * math.pi)))
    for m in range(lmax + 1):
        if m > 0:
            pmm = -pmm * math.sqrt((2.0 * m + 1.0) / (2.0 * m)) * s
        P[m, m] = pmm
        if m + 1 <= lmax:
            P[m + 1, m] = x * math.sqrt(2.0 * m + 3.0) * pmm
        for l in range(m + 2, lmax + 1):
            a = math.sqrt((4.0 * l * l - 1.0) / (l * l - m * m))
            b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1.0) ** 2 - 1.0))
            P[l, m] = a * (x * P[l - 1, m] - b * P[l - 2, m])
    return P


def sh_noise(n_theta: int, n_phi: int, lmax: int, seed: int) -> np.ndarray:
    """Real spherical-harmonic series of degree <= lmax with seeded Gaussian
    coefficients (weight 1/(l+1)), scaled to a peak of 1."""
    rng = np.random.default_rng(seed)
    theta, phi = grid_angles(n_theta, n_phi)
    P = _legendre_normalized(lmax, np.cos(theta))
    A = np.zeros((n_theta, lmax + 1))
    B = np.zeros((n_theta, lmax + 1))
    for l in range(lmax + 1):
        w = 1.0 / (l + 1.0)
        a = rng.standard_normal(l + 1) * w
        b = rng.standard_normal(l + 1) * w
        b[0] = 0.0
        A[:, : l + 1] += P[l, : l + 1].T * a
        B[:, : l + 1] += P[l, : l + 1].T * b
    m = np.arange(lmax + 1)
    cosm = np.cos(np.outer(m, phi))
    sinm = np.sin(np.outer(m, phi))
    f = A @ cosm + B @ sinm
    return f / np.max(np.abs(f))


def smooth_bubble(n_theta: int, n_phi: int, seed: int, lmax: int = 16, amp: float = 0.10,
                  hmin: float = 200e-9, hmax: float = 1000e-9) -> np.ndarray:
    rng = np.random.default_rng(seed + 7919)
    base = rng.uniform(hmin, hmax)
    return ff_exact(base * (1.0 + amp * sh_noise(n_theta, n_phi, lmax, seed)))


def round_bf16(x) -> np.ndarray:
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    b = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    return b.astype(np.uint32).view(np.float32).astype(np.float64)


def spectral_cmf(seed: int, bf16: bool = True) -> np.ndarray:
    """cmf[k, c] for the 81 wavelengths and three seeded Gaussian curves,
    stored at BF16 precision (the MFMA operand type) unless bf16 is False."""
    rng = np.random.default_rng(seed)
    centres = np.array([610.0, 550.0, 460.0]) + rng.uniform(-15.0, 15.0, 3)
    widths = rng.uniform(25.0, 45.0, 3)
    heights = rng.uniform(0.8, 1.2, 3)
    lam = WAVELENGTHS_NM[:, None]
    cmf = heights * np.exp(-0.5 * ((lam - centres) / widths) ** 2)
    return round_bf16(cmf) if bf16 else cmf


def camera(radius: float) -> Tuple[float, float, float]:
    """Fixed synthetic camera, 4 bubble radii out, above the equator."""
    return (2.2 * radius, 1.4 * radius, 3.1 * radius)


def foam_batch(n_bubbles: int, seed: int, lat=(32, 512), diameter=(5e-3, 50e-3),
               h=(200e-9, 1000e-9), viscosity=(1e-3, 5e-3)):
    """Bubble descriptions of a batched foam: n_theta ~ logU[lat] (rounded),
    n_phi = 2 n_theta, diameter ~ logU, base thickness and viscosity ~ U.
    Returns a list of dicts with n_theta, n_phi, bubble_radius, viscosity, h0."""
    rng = np.random.default_rng(seed)
    lat_v = np.exp(rng.uniform(math.log(lat[0]), math.log(lat[1]), n_bubbles))
    nt = np.clip(np.rint(lat_v).astype(np.int64), lat[0], lat[1])
    d = np.exp(rng.uniform(math.log(diameter[0]), math.log(diameter[1]), n_bubbles))
    h0 = rng.uniform(h[0], h[1], n_bubbles)
    eta = rng.uniform(viscosity[0], viscosity[1], n_bubbles)
    return [dict(n_theta=int(nt[k]), n_phi=int(2 * nt[k]), bubble_radius=float(d[k] / 2.0),
                 viscosity=float(eta[k]), h0=float(h0[k])) for k in range(n_bubbles)]


def foam_noise(offsets: np.ndarray, sizes, h0, seed: int, rel: float = 0.05,
               dtype=np.float64) -> np.ndarray:
    """Concatenated thickness fields h0_b (1 + rel U[-1, 1]) per cell; padding
    cells between bubbles hold h0 of the bubble before them."""
    rng = np.random.default_rng(seed)
    total = int(offsets[-1])
    out = np.empty(total, dtype=dtype)
    for k in range(len(sizes)):
        a, b = int(offsets[k]), int(offsets[k + 1])
        n = int(sizes[k])
        out[a:a + n] = h0[k] * (1.0 + rel * rng.uniform(-1.0, 1.0, n))
        out[a + n:b] = h0[k]
    return out


def cluster_bubbles(n: int, seed: int, box: float = 0.25, radius=(2.5e-3, 25e-3), speed: float = 0.02):
    """Seeded foam cluster: positions uniform in a cube of side ``box`` (m),
    radii log-uniform in ``radius``, velocity components uniform in
    [-speed, speed] m/s, film thickness U[200, 1000] nm; FP32 values.
    Returns (pos [n][3], radius [n], vel [n][3], thickness_nm [n])."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-box / 2.0, box / 2.0, (n, 3)).astype(np.float32)
    r = np.exp(rng.uniform(math.log(radius[0]), math.log(radius[1]), n)).astype(np.float32)
    vel = rng.uniform(-speed, speed, (n, 3)).astype(np.float32)
    th = rng.uniform(200.0, 1000.0, n).astype(np.float32)
    return pos, r, vel, th
