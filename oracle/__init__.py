"""ctypes view of the oracle library -- TEST INFRASTRUCTURE ONLY.

The oracle is a CPU restatement of the reference RSF iteration
(CUDA_RSF/main.cpp:222-300 and the kernels of CUDA_RSF/kernel.cu).  Only the
test suite, the smoke check and the CPU-baseline leg of the benchmark import
this module, as the checker or as the timed CPU baseline.  The product
(paper-rsf-vesselseg_amd/) never imports it.

Arrays are numpy volumes of shape (l, h, w) = (Z, Y, X), C order, so that the
reference's X-fastest index ((z*h)+y)*w+x is the flat index.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class Params(C.Structure):
    _fields_ = [
        ("sigma", C.c_float),
        ("lambda1", C.c_float),
        ("lambda2", C.c_float),
        ("nu", C.c_float),
        ("mu", C.c_float),
        ("dt", C.c_float),
        ("epsilon", C.c_float),
        ("sigma2", C.c_float),
    ]


def make_params(sigma=3.0, lambda1=1.0, lambda2=None, nu=65.025, mu=1.0, dt=0.1,
                epsilon=0.2, sigma2=0.0) -> Params:
    if lambda2 is None:
        lambda2 = lambda1
    return Params(sigma, lambda1, lambda2, nu, mu, dt, epsilon, sigma2)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (g++ -O3 -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        i64 = C.c_int64
        ci = C.c_int
        cf = C.c_float
        sig = {
            "oracle_taps": (ci, [cf, f32p, ci, C.POINTER(ci), C.POINTER(ci)]),
            "oracle_atan_f64": (None, [f64p, f64p, i64]),
            "oracle_heaviside": (None, [f32p, f32p, i64, cf]),
            "oracle_deri_heaviside": (None, [f32p, f32p, i64, cf]),
            "oracle_multiplication": (None, [f32p, f32p, f32p, i64]),
            "oracle_inverse": (None, [f32p, f32p, i64]),
            "oracle_plus": (None, [f32p, f32p, f32p, f32p, i64]),
            "oracle_division": (None, [f32p, f32p, f32p, i64]),
            "oracle_division_grad": (None, [f32p, f32p, f32p, i64]),
            "oracle_gradientmagnitude": (None, [f32p, f32p, f32p, f32p, i64]),
            "oracle_gradient": (None, [f32p, f32p, f32p, f32p, ci, ci, ci]),
            "oracle_conv_separable": (None, [f32p, f32p, ci, ci, ci, f32p, ci]),
            "oracle_pad_border": (None, [f32p, f32p, ci, ci, ci, ci]),
            "oracle_energies": (None, [f32p, f32p, f32p, f32p, f32p, ci, ci, ci, ci]),
            "oracle_final_phi": (None, [f32p, f32p, f32p, f32p, f32p, f32p, f32p,
                                        cf, cf, cf, cf, cf, i64]),
            "oracle_step": (ci, [f32p, f32p, ci, ci, ci, C.POINTER(Params)]),
            "oracle_run": (ci, [f32p, f32p, ci, ci, ci, C.POINTER(Params), ci, ci]),
            "oracle_step_debug": (ci, [f32p, f32p, ci, ci, ci, C.POINTER(Params)] + [C.c_void_p] * 7),
            "oracle_step_f64": (ci, [f64p, f64p, ci, ci, ci, C.POINTER(Params)]),
            "oracle_step_adjoint_f64": (ci, [f64p, f64p, ci, ci, ci, C.POINTER(Params), f64p, f64p,
                                             f64p]),
            "oracle_diff_f64": (None, [f64p, f64p, ci, ci, ci, ci]),
            "oracle_diff_transpose_f64": (None, [f64p, f64p, ci, ci, ci, ci]),
            "oracle_max_threads": (ci, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dims(vol):
    l, h, w = vol.shape
    return w, h, l


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def taps(sigma: float):
    g = np.zeros(512, np.float32)
    P, m = C.c_int(), C.c_int()
    rc = lib().oracle_taps(sigma, g, 512, C.byref(P), C.byref(m))
    if rc:
        raise ValueError(f"sigma={sigma}: unsupported (floor(7*sigma) even or too small)")
    return g[: P.value].copy(), P.value, m.value


def atan_f64(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().oracle_atan_f64(out, x, x.size)
    return out


def heaviside(phi, eps=0.2):
    phi = _f32(phi)
    out = np.empty_like(phi)
    lib().oracle_heaviside(out, phi, phi.size, eps)
    return out


def deri_heaviside(phi, eps=0.2):
    phi = _f32(phi)
    out = np.empty_like(phi)
    lib().oracle_deri_heaviside(out, phi, phi.size, eps)
    return out


def gradient(f):
    f = _f32(f)
    gx, gy, gz = (np.empty_like(f) for _ in range(3))
    lib().oracle_gradient(gx, gy, gz, f, *_dims(f))
    return gx, gy, gz


def gradientmagnitude(gx, gy, gz):
    gx, gy, gz = _f32(gx), _f32(gy), _f32(gz)
    out = np.empty_like(gx)
    lib().oracle_gradientmagnitude(out, gx, gy, gz, gx.size)
    return out


def division_grad(g, M):
    g, M = _f32(g), _f32(M)
    out = np.empty_like(g)
    lib().oracle_division_grad(out, g, M, g.size)
    return out


def elementwise(name, *arrs):
    arrs = [_f32(a) for a in arrs]
    out = np.empty_like(arrs[0])
    getattr(lib(), "oracle_" + name)(out, *arrs, arrs[0].size)
    return out


def conv_separable(vol, g):
    vol = _f32(vol)
    g = _f32(g)
    P = g.size
    l, h, w = vol.shape
    out = np.empty((l - P + 1, h - P + 1, w - P + 1), np.float32)
    lib().oracle_conv_separable(out, vol, w, h, l, g, P)
    return out


def pad_border(valid, pad):
    valid = _f32(valid)
    vl, vh, vw = valid.shape
    out = np.empty((vl + 2 * pad, vh + 2 * pad, vw + 2 * pad), np.float32)
    lib().oracle_pad_border(out, valid, vw, vh, vl, pad)
    return out


def energies(fo, fi, image, s=0):
    fo, fi, image = _f32(fo), _f32(fi), _f32(image)
    Eo, Ei = np.empty_like(fo), np.empty_like(fo)
    lib().oracle_energies(Eo, Ei, fo, fi, image, *_dims(image), int(s))
    return Eo, Ei


def final_phi(phi, dh, Eo, Ei, kappa, L, lambda1, lambda2, dt, nu, mu):
    a = [_f32(x) for x in (phi, dh, Eo, Ei, kappa, L)]
    out = np.empty_like(a[0])
    lib().oracle_final_phi(out, *a, lambda1, lambda2, dt, nu, mu, a[0].size)
    return out


def step(phi, image, params: Params):
    phi = _f32(phi).copy()
    rc = lib().oracle_step(phi, _f32(image), *_dims(phi), C.byref(params))
    if rc:
        raise ValueError(f"oracle_step failed ({rc})")
    return phi


def run(phi, image, params: Params, T: int, threads: int = 0):
    phi = _f32(phi).copy()
    rc = lib().oracle_run(phi, _f32(image), *_dims(phi), C.byref(params), int(T), int(threads))
    if rc:
        raise ValueError(f"oracle_run failed ({rc})")
    return phi


def step_debug(phi, image, params: Params):
    phi = _f32(phi)
    outs = {k: np.empty_like(phi) for k in ("phi", "H", "F", "kappa", "L", "fo", "fi")}
    ptrs = [outs[k].ctypes.data_as(C.c_void_p) for k in ("phi", "H", "F", "kappa", "L", "fo", "fi")]
    rc = lib().oracle_step_debug(phi, _f32(image), *_dims(phi), C.byref(params), *ptrs)
    if rc:
        raise ValueError(f"oracle_step_debug failed ({rc})")
    return outs


def step_f64(phi, image, params: Params):
    phi = np.ascontiguousarray(phi, np.float64).copy()
    rc = lib().oracle_step_f64(phi, np.ascontiguousarray(image, np.float64), *_dims(phi),
                               C.byref(params))
    if rc:
        raise ValueError(f"oracle_step_f64 failed ({rc})")
    return phi


def step_adjoint_f64(phi, image, params: Params, gout):
    phi = np.ascontiguousarray(phi, np.float64)
    gin = np.zeros_like(phi)
    gp = np.zeros(4, np.float64)
    rc = lib().oracle_step_adjoint_f64(phi, np.ascontiguousarray(image, np.float64), *_dims(phi),
                                       C.byref(params), np.ascontiguousarray(gout, np.float64),
                                       gin, gp)
    if rc:
        raise ValueError(f"oracle_step_adjoint_f64 failed ({rc})")
    return gin, gp


def diff_f64(f, axis):
    f = np.ascontiguousarray(f, np.float64)
    out = np.empty_like(f)
    lib().oracle_diff_f64(out, f, *_dims(f), int(axis))
    return out


def diff_transpose_f64(u, axis):
    u = np.ascontiguousarray(u, np.float64)
    out = np.empty_like(u)
    lib().oracle_diff_transpose_f64(out, u, *_dims(u), int(axis))
    return out


def max_threads() -> int:
    return int(lib().oracle_max_threads())
