"""FP64 numpy restatement of the Fourier neural operator surrogate.

TEST INFRASTRUCTURE ONLY.  The surrogate has no counterpart in the
reference; this module is the checker of the HIP forward and backward
kernels (``sf_fno_forward`` / ``sf_fno_gradient``).  It is formulated
independently of the HIP path: spectral convolutions use ``numpy.fft``
(rfft2 / irfft2 and their adjoints via fft2 / ifft2) where the HIP path
multiplies by explicit truncated DFT matrices, and the backward pass is
written out by hand.  Its gradient is pinned by FP64 central differences
in ``tests/test_fno.py``.

Parameter layout (flat, as ``sf_fno_get_params``): lift W[8][C], b[C];
per layer W[C][C], b[C], R[2 Mt'][Mp][2][C][C] (re, im; theta modes
k = 0..Mt'-1, -Mt'..-1); projection Q[C][2], q[2].
"""
from __future__ import annotations

import numpy as np

N_IN = 8
N_OUT = 2
K0 = 0.7978845608028654
K1 = 0.044715


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(K0 * (z + K1 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(K0 * (z + K1 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * z * z)


class Shapes:
    def __init__(self, n_theta, n_phi, width, modes_theta, modes_phi, layers):
        self.T, self.P, self.C = n_theta, n_phi, width
        self.Mt, self.Mp, self.L = 2 * modes_theta, modes_phi, layers
        self.kidx = np.array([k if k < modes_theta else k - 2 * modes_theta for k in range(2 * modes_theta)]) % n_theta

    @classmethod
    def of(cls, cfg):
        return cls(cfg.n_theta, cfg.n_phi, cfg.width, cfg.modes_theta, cfg.modes_phi, cfg.layers)

    def n_params(self):
        C, L = self.C, self.L
        return N_IN * C + C + L * (C * C + C + self.Mt * self.Mp * 2 * C * C) + C * N_OUT + N_OUT


def unpack(flat, s: Shapes):
    flat = np.asarray(flat, np.float64)
    o = 0

    def take(n, shape):
        nonlocal o
        a = flat[o:o + n].reshape(shape)
        o += n
        return a
    C = s.C
    p = {"Wl": take(N_IN * C, (N_IN, C)), "bl": take(C, (C,)), "W": [], "b": [], "R": []}
    for _ in range(s.L):
        p["W"].append(take(C * C, (C, C)))
        p["b"].append(take(C, (C,)))
        r = take(s.Mt * s.Mp * 2 * C * C, (s.Mt, s.Mp, 2, C, C))
        p["R"].append(r[:, :, 0] + 1j * r[:, :, 1])
    p["Q"] = take(C * N_OUT, (C, N_OUT))
    p["q"] = take(N_OUT, (N_OUT,))
    assert o == flat.size
    return p


def pack_grad(g, s: Shapes):
    parts = [g["Wl"].ravel(), g["bl"].ravel()]
    for l in range(s.L):
        r = np.stack([g["R"][l].real, g["R"][l].imag], axis=2)
        parts += [g["W"][l].ravel(), g["b"][l].ravel(), r.ravel()]
    parts += [g["Q"].ravel(), g["q"].ravel()]
    return np.concatenate(parts)


def spectral(x, R, s: Shapes):
    """x [B,T,P,C] -> (y [B,T,P,C], X [B,Mt,Mp,C] complex)."""
    F = np.fft.rfft2(x, axes=(1, 2))                      # [B,T,P/2+1,C]
    X = F[:, s.kidx][:, :, :s.Mp]                         # [B,Mt,Mp,C]
    Y = np.einsum("bkmc,kmco->bkmo", X, R)
    full = np.zeros((x.shape[0], s.T, s.P // 2 + 1, s.C), complex)
    full[:, s.kidx, :s.Mp] = Y
    y = np.fft.irfft2(full, s=(s.T, s.P), axes=(1, 2))
    return y, X


def spectral_adjoint(dy, X, R, s: Shapes):
    """Adjoint of `spectral` at fixed R: (dx, dR)."""
    T, P = s.T, s.P
    w = np.where(np.arange(s.Mp) == 0, 1.0, 2.0) / (T * P)
    dF = np.fft.fft2(dy, axes=(1, 2))[:, s.kidx][:, :, :s.Mp]
    dY = dF * w[None, None, :, None]
    dR = np.einsum("bkmc,bkmo->kmco", np.conj(X), dY)
    dX = np.einsum("bkmo,kmco->bkmc", dY, np.conj(R))
    full = np.zeros((dy.shape[0], T, P, s.C), complex)
    full[:, s.kidx, :s.Mp] = dX
    dx = np.real(np.fft.ifft2(full, axes=(1, 2))) * (T * P)
    return dx, dR


def forward(flat, x_in, s: Shapes):
    """x_in [B,T,P,N_IN] -> out [B,T,P,2] and the cache of the backward pass."""
    p = unpack(flat, s)
    x = np.asarray(x_in, np.float64) @ p["Wl"] + p["bl"]
    cache = {"x": [x], "z": [], "X": []}
    for l in range(s.L):
        y, X = spectral(x, p["R"][l], s)
        z = x @ p["W"][l] + p["b"][l] + y
        x = gelu(z) if l < s.L - 1 else z
        cache["z"].append(z)
        cache["X"].append(X)
        cache["x"].append(x)
    out = x @ p["Q"] + p["q"]
    return out, cache, p


def loss_and_grad(flat, x_in, target, s: Shapes):
    out, cache, p = forward(flat, x_in, s)
    target = np.asarray(target, np.float64)
    n = out.size
    loss = float(np.sum((out - target) ** 2) / n)
    dout = 2.0 * (out - target) / n
    g = {"W": [None] * s.L, "b": [None] * s.L, "R": [None] * s.L}
    xL = cache["x"][s.L]
    g["Q"] = np.einsum("btpc,btpj->cj", xL, dout)
    g["q"] = dout.sum(axis=(0, 1, 2))
    dx = dout @ p["Q"].T
    for l in range(s.L - 1, -1, -1):
        z = cache["z"][l]
        dz = dx * gelu_grad(z) if l < s.L - 1 else dx
        g["b"][l] = dz.sum(axis=(0, 1, 2))
        g["W"][l] = np.einsum("btpc,btpo->co", cache["x"][l], dz)
        dxs, dR = spectral_adjoint(dz, cache["X"][l], p["R"][l], s)
        g["R"][l] = dR
        dx = dz @ p["W"][l].T + dxs
    x_in = np.asarray(x_in, np.float64)
    g["Wl"] = np.einsum("btpi,btpc->ic", x_in, dx)
    g["bl"] = dx.sum(axis=(0, 1, 2))
    return loss, pack_grad(g, s), out
