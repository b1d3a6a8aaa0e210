"""GPU parity of the reference-wrapper mirrors (host pointers, C ABI) against
the oracle, one reference function at a time (CUDA_RSF/kernel.cu)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rsf_amd

pytestmark = pytest.mark.gpu

SHAPE = (13, 17, 19)  # (l, h, w): odd, non-cubic, not a multiple of any tile


def _p(a):
    return a.ctypes.data


def _rand(seed, lo=-5, hi=5, shape=SHAPE):
    return np.random.default_rng(seed).uniform(lo, hi, shape).astype(np.float32)


def _dims(shape=SHAPE):
    l, h, w = shape
    return w, h, l


def _ulps(a, b):
    a = a.astype(np.float32).view(np.int32).astype(np.int64)
    b = b.astype(np.float32).view(np.int32).astype(np.int64)
    return int(np.abs(a - b).max())


def test_mirror_heaviside_and_dirac():
    phi = _rand(1, -3, 3)
    phi.ravel()[:3] = [0.0, 10.0, -10.0]
    out = np.empty_like(phi)
    rsf_amd.check(rsf_amd.lib().rsf_heaviside_cal(_p(out), _p(phi), *_dims()))
    # both sides evaluate the same correctly rounded atanf: bit-identical
    assert np.array_equal(out, oracle.heaviside(phi))
    rsf_amd.check(rsf_amd.lib().rsf_deri_heaviside_cal(_p(out), _p(phi), *_dims()))
    assert np.array_equal(out, oracle.deri_heaviside(phi))


def test_mirror_elementwise():
    a, b, c = _rand(2), _rand(3, 0.5, 4), _rand(4)
    out = np.empty_like(a)
    L = rsf_amd.lib()
    rsf_amd.check(L.rsf_inverse_cal(_p(out), _p(a), *_dims()))
    assert np.array_equal(out, oracle.elementwise("inverse", a))
    rsf_amd.check(L.rsf_multiplication_cal(_p(out), _p(a), _p(b), *_dims()))
    assert np.array_equal(out, a * b)
    rsf_amd.check(L.rsf_division_cal(_p(out), _p(a), _p(b), *_dims()))
    assert np.array_equal(out, a / b)
    rsf_amd.check(L.rsf_normalized_div_plus(_p(out), _p(a), _p(b), _p(c), *_dims()))
    assert np.array_equal(out, (a + b) + c)
    rsf_amd.check(L.rsf_gradientmagnitude_cal(_p(out), _p(a), _p(b), _p(c), *_dims()))
    assert np.array_equal(out, oracle.gradientmagnitude(a, b, c))


def test_mirror_gradients_with_wrap():
    f = _rand(5)
    gx, gy, gz = (np.empty_like(f) for _ in range(3))
    rsf_amd.check(rsf_amd.lib().rsf_gradient_cal(_p(f), _p(gx), _p(gy), _p(gz), *_dims()))
    ox, oy, oz = oracle.gradient(f)
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy) and np.array_equal(gz, oz)


def test_mirror_normals_curvature_laplacian():
    f = _rand(6)
    gx, gy, gz = oracle.gradient(f)
    M = oracle.gradientmagnitude(gx, gy, gz)
    nx, ny, nz = (np.empty_like(f) for _ in range(3))
    L = rsf_amd.lib()
    rsf_amd.check(L.rsf_division_nDIV(_p(gx), _p(gy), _p(gz), _p(M), _p(nx), _p(ny), _p(nz), *_dims()))
    assert np.array_equal(nx, oracle.division_grad(gx, M))
    assert np.array_equal(nz, oracle.division_grad(gz, M))
    dx, dy, dz = (np.empty_like(f) for _ in range(3))
    rsf_amd.check(L.rsf_gradient_cal_3_input(_p(nx), _p(ny), _p(nz), _p(dx), _p(dy), _p(dz), *_dims()))
    assert np.array_equal(dx, oracle.gradient(nx)[0])
    assert np.array_equal(dy, oracle.gradient(ny)[1])
    assert np.array_equal(dz, oracle.gradient(nz)[2])
    div = np.empty_like(f)
    rsf_amd.check(L.rsf_DIV_cal(_p(div), _p(gx), _p(gy), _p(gz), *_dims()))
    ref = oracle.elementwise("plus", oracle.gradient(gx)[0], oracle.gradient(gy)[1], oracle.gradient(gz)[2])
    assert np.array_equal(div, ref)


@pytest.mark.parametrize("sigma,shape", [(1.0, (13, 17, 19)), (3.0, (23, 25, 27))])
def test_mirror_separable_convolution(sigma, shape):
    g, P, m = oracle.taps(sigma)
    v = _rand(7, 0, 255, shape)
    l, h, w = shape
    out = np.empty((l - P + 1, h - P + 1, w - P + 1), np.float32)
    rsf_amd.check(rsf_amd.lib().rsf_adddevice_convolution_seperable(_p(out), _p(v), w, h, l, sigma, _p(g), P))
    assert np.array_equal(out, oracle.conv_separable(v, g))


def test_mirror_pad_border():
    v = _rand(8, 0, 1, (5, 6, 7))
    out = np.empty((9, 10, 11), np.float32)
    rsf_amd.check(rsf_amd.lib().rsf_pad_border_cal(_p(out), _p(v), 7, 6, 5, 2))
    assert np.array_equal(out, oracle.pad_border(v, 2))


@pytest.mark.parametrize("s", [0.0, 1.0, 1.7])
def test_mirror_energies(s):
    I = _rand(9, 0, 255)
    fo, fi = _rand(10, 0, 255), _rand(11, 0, 255)
    Eo, Ei = np.empty_like(I), np.empty_like(I)
    rsf_amd.check(rsf_amd.lib().rsf_adddevice(_p(I), _p(fo), _p(fi), _p(Eo), _p(Ei), *_dims(), s))
    ro, ri = oracle.energies(fo, fi, I, int(s))
    assert np.array_equal(Eo, ro) and np.array_equal(Ei, ri)


def test_mirror_final_phi():
    a = [_rand(k) for k in range(12, 18)]
    out = np.empty_like(a[0])
    rsf_amd.check(rsf_amd.lib().rsf_Final_PHI_cal(_p(out), *[_p(x) for x in a], 0.1, 0.1, 1e-4, 1.0, *_dims()))
    assert np.array_equal(out, oracle.final_phi(*a, 0.1, 0.1, 0.1, 1e-4, 1.0))


def test_mirror_rejects_bad_arguments():
    a = _rand(1)
    with pytest.raises(rsf_amd.RSFError):
        rsf_amd.check(rsf_amd.lib().rsf_heaviside_cal(_p(a), None, *_dims()))
    with pytest.raises(rsf_amd.RSFError):
        rsf_amd.check(rsf_amd.lib().rsf_heaviside_cal(_p(a), _p(a), 0, 3, 3))


def test_mirror_heaviside_wide_range():
    """The device arctangent (degree-9 polynomial, unscaled division) on ~4M
    inputs spread over magnitudes 2^-30 .. 2^30, and the runs of consecutive
    floats where the arctangent is flattest, against the oracle bit for bit."""
    rng = np.random.default_rng(7)
    shape = (64, 256, 256)
    n = int(np.prod(shape))
    mag = np.exp2(rng.uniform(-30, 30, n))
    phi = (mag * rng.choice([-1.0, 1.0], n)).astype(np.float32)
    runs = [np.float32(62921456.0) * 0.2, np.float32(0.00035211083) * 0.2, np.float32(1.0278158) * 0.2]
    k = 0
    for c in runs:
        b = np.array([c], np.float32).view(np.int32)[0]
        phi[k:k + 4096] = (b + np.arange(-2048, 2048, dtype=np.int32)).view(np.float32)
        k += 4096
    phi[k:k + 6] = [0.0, -0.0, np.inf, -np.inf, 1e-45, -3e38]
    phi = phi.reshape(shape)
    out = np.empty_like(phi)
    rsf_amd.check(rsf_amd.lib().rsf_heaviside_cal(_p(out), _p(phi), *_dims(shape)))
    assert np.array_equal(out, oracle.heaviside(phi))
