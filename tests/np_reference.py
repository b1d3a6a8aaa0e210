"""Independent FP64 numpy restatement of one reference RSF iteration.

Written separately from the C++ oracle, in a different formulation, so that the
two cross-check each other on small volumes:

* differences use linear-index neighbour tables (the reference's x = 0 wrap,
  CUDA_RSF/kernel.cu:64-74, is simply "previous linear element");
* the valid separable correlation uses sliding windows (kernel.cu:1000-1049);
* the whole iteration follows CUDA_RSF/main.cpp:234-300.

Arrays have shape (Z, Y, X) = (l, h, w).
"""
from __future__ import annotations

import numpy as np
from numpy.lib.stride_tricks import sliding_window_view


def taps64(sigma: float):
    """Closed form of CUDA_RSF/main.cpp:152-166 evaluated in float64."""
    p = int(np.float32(7) * np.float32(sigma))
    if p % 2 == 0:
        raise ValueError("even tap count")
    miu = p // 2
    u = int(np.float32(2) * np.float32(sigma) * np.float32(sigma))
    k = np.arange(p, dtype=np.float64)
    g = 1.0 / np.sqrt(u * np.pi) * np.exp(-((k - miu) ** 2) / u)
    m = (p - 1) // 2
    return g, p, m


def diff(f: np.ndarray, axis: int) -> np.ndarray:
    """Half central difference with the reference's edge rules; axis 0=x,1=y,2=z."""
    l, h, w = f.shape
    if axis == 0:
        flat = f.ravel()
        idx = np.arange(flat.size)
        right = np.where(idx % w == w - 1, idx, idx + 1)
        left = np.maximum(idx - 1, 0)
        return ((flat[right] - flat[left]) / 2.0).reshape(f.shape)
    ax = 1 if axis == 1 else 0
    out = np.empty_like(f)
    sl = [slice(None)] * 3

    def s(a, b):
        t = list(sl)
        t[ax] = slice(a, b)
        return tuple(t)

    out[s(1, -1)] = (f[s(2, None)] - f[s(None, -2)]) / 2.0
    out[s(0, 1)] = (f[s(1, 2)] - f[s(0, 1)]) / 2.0
    out[s(-1, None)] = (f[s(-1, None)] - f[s(-2, -1)]) / 2.0
    return out


def blur_valid(v: np.ndarray, g: np.ndarray) -> np.ndarray:
    P = g.size
    for ax in (2, 1, 0):  # x, then y, then z
        v = np.tensordot(sliding_window_view(v, P, axis=ax), g, axes=([3], [0]))
    return v


def pad30(valid: np.ndarray, m: int) -> np.ndarray:
    out = np.full(tuple(s + 2 * m for s in valid.shape), 30.0)
    out[m:-m or None, m:-m or None, m:-m or None] = valid
    return out


def energies(I, fo, fi, s: int):
    if s == 0:
        return (I - fo) ** 2, (I - fi) ** 2
    l, h, w = I.shape
    Eo = np.zeros_like(I)
    Ei = np.zeros_like(I)
    wgt = 1.0 / (2 * s + 1) ** 2
    for dz in range(-s, s + 1):
        for dy in range(-s, s + 1):
            for dx in range(-s, s + 1):
                sh = np.full_like(I, np.nan)
                zs = slice(max(0, -dz), min(l, l - dz))
                ys = slice(max(0, -dy), min(h, h - dy))
                xs = slice(max(0, -dx), min(w, w - dx))
                zt = slice(max(0, dz), min(l, l + dz))
                yt = slice(max(0, dy), min(h, h + dy))
                xt = slice(max(0, dx), min(w, w + dx))
                sh[zs, ys, xs] = I[zt, yt, xt]
                ok = ~np.isnan(sh)
                Eo[ok] += wgt * (sh[ok] - fo[ok]) ** 2
                Ei[ok] += wgt * (sh[ok] - fi[ok]) ** 2
    return Eo, Ei


def step(phi, I, sigma=3.0, lambda1=1.0, lambda2=None, nu=65.025, mu=1.0, dt=0.1, eps=0.2,
         sigma2=0.0, g=None, debug=False):
    """One FP64 iteration; `g` overrides the taps (e.g. the fp32 taps as float64)."""
    lambda2 = lambda1 if lambda2 is None else lambda2
    phi = np.asarray(phi, np.float64)
    I = np.asarray(I, np.float64)
    if g is None:
        g, P, m = taps64(sigma)
    else:
        P = g.size
        m = (P - 1) // 2
    H = 0.5 * (1 + (2 / np.pi) * np.arctan(phi / eps))
    dh = (1 / np.pi) * (eps / (eps * eps + phi * phi))
    gx, gy, gz = diff(phi, 0), diff(phi, 1), diff(phi, 2)
    M = np.sqrt(gx * gx + gy * gy + gz * gz)
    nx, ny, nz = gx / (M + 1e-6), gy / (M + 1e-6), gz / (M + 1e-6)
    kappa = diff(nx, 0) + diff(ny, 1) + diff(nz, 2)
    L = diff(gx, 0) + diff(gy, 1) + diff(gz, 2)
    fo = pad30(blur_valid(I * H, g), m) / pad30(blur_valid(H, g), m)
    fi = pad30(blur_valid(I * (1 - H), g), m) / pad30(blur_valid(1 - H, g), m)
    Eo, Ei = energies(I, fo, fi, int(sigma2))
    F = lambda1 * Eo - lambda2 * Ei
    out = phi - dh * F * dt + nu * dt * dh * kappa + mu * dt * (L - kappa)
    if debug:
        return dict(phi=out, H=H, F=F, kappa=kappa, L=L, fo=fo, fi=fi)
    return out
