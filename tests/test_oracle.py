"""Pins the oracle (CPU restatement of the reference) against closed forms,
known answers, brute force and an independent FP64 numpy restatement.

Each test is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign or neighbour index, x/y transposed, the wrap missing) fails it.
"""
import numpy as np
import pytest

import np_reference as npr
import oracle
from conftest import small_scene


# ----------------------------------------------------------------- taps (main.cpp:152-166)
def test_oracle_taps_closed_form():
    g, P, m = oracle.taps(3.0)
    assert (P, m) == (21, 10)
    u, miu = 18, 10
    k = np.arange(21)
    closed = 1 / np.sqrt(u * np.pi) * np.exp(-((k - miu) ** 2) / u)
    np.testing.assert_allclose(g, closed, rtol=3e-7)
    assert abs(g[10] - 0.132980749) < 1e-8
    assert abs(float(g.sum(dtype=np.float64)) - 0.999562079) < 2e-7
    assert np.array_equal(g, g[::-1])


@pytest.mark.parametrize("sigma,P,m,u", [(1.0, 7, 3, 2), (5.0, 35, 17, 50), (2.5, 17, 8, 12),
                                         (3.0, 21, 10, 18)])
def test_oracle_tap_integer_rules(sigma, P, m, u):
    g, P2, m2 = oracle.taps(sigma)
    assert (P2, m2) == (P, m)
    miu = P // 2
    k = np.arange(P)
    np.testing.assert_allclose(g, 1 / np.sqrt(u * np.pi) * np.exp(-((k - miu) ** 2) / u), rtol=3e-7)


@pytest.mark.parametrize("sigma", [2.0, 4.0, 0.1])
def test_oracle_taps_reject_even(sigma):
    with pytest.raises(ValueError):
        oracle.taps(sigma)


# ---------------------------------------------------- Heaviside / Dirac (kernel.cu:248-280)
def test_oracle_heaviside_known_answers():
    H = oracle.heaviside(np.array([0.0, 10.0, -10.0], np.float32))
    assert H[0] == np.float32(0.5)
    assert abs(H[1] - 0.993634641) < 1e-8
    assert abs(H[2] - 0.006365343) < 1e-8
    x = np.linspace(-30, 30, 2001).astype(np.float32)
    Hp, Hm = oracle.heaviside(x), oracle.heaviside(-x)
    np.testing.assert_allclose(Hp + Hm, 1.0, atol=1.2e-7)


def test_oracle_dirac_known_answers_and_derivative():
    d = oracle.deri_heaviside(np.array([0.0, 10.0], np.float32))
    assert abs(d[0] - 1.59154928) < 2e-7
    assert abs(d[1] - 6.36365206e-4) < 1e-11
    # delta is dH/dphi: check against a float64 central difference of H
    x = np.linspace(-3, 3, 61)
    h = 1e-5
    dH = ((0.5 * (1 + 2 / np.pi * np.arctan((x + h) / 0.2))) -
          (0.5 * (1 + 2 / np.pi * np.arctan((x - h) / 0.2)))) / (2 * h)
    np.testing.assert_allclose(oracle.deri_heaviside(x.astype(np.float32)), dH, rtol=1e-5)


def test_oracle_epsilon_parameter():
    x = np.array([0.5, -2.0], np.float32)
    np.testing.assert_allclose(oracle.heaviside(x, 1.0),
                               0.5 * (1 + 2 / np.pi * np.arctan(x.astype(np.float64))), rtol=1e-6)


# -------------------------------------------------------- element-wise (kernel.cu:173-298)
def test_oracle_elementwise_identities():
    rng = np.random.default_rng(1)
    a, b, c = (rng.normal(size=(4, 5, 6)).astype(np.float32) for _ in range(3))
    assert np.array_equal(oracle.elementwise("multiplication", a, b), a * b)
    assert np.array_equal(oracle.elementwise("inverse", a), np.float32(1) - a)
    assert np.array_equal(oracle.elementwise("plus", a, b, c), (a + b) + c)
    assert np.array_equal(oracle.elementwise("division", a, b), a / b)
    M = np.abs(b)
    ref = (a.astype(np.float64) / (M.astype(np.float64) + 1e-6)).astype(np.float32)
    assert np.array_equal(oracle.division_grad(a, M), ref)
    ref = np.sqrt((a * a + b * b) + c * c)
    assert np.array_equal(oracle.gradientmagnitude(a, b, c), ref)


# ---------------------------------------------------------- gradients (kernel.cu:57-171)
def test_oracle_gradient_ramp_closed_forms():
    l, h, w = 6, 7, 8
    a, b, c = 0.5, -1.25, 2.0
    z, y, x = np.meshgrid(np.arange(l), np.arange(h), np.arange(w), indexing="ij")
    f = (a * x + b * y + c * z).astype(np.float32)
    gx, gy, gz = oracle.gradient(f)
    # interior
    assert np.allclose(gx[:, :, 1:-1], a) and np.allclose(gy[:, 1:-1, :], b)
    assert np.allclose(gz[1:-1], c)
    # one-sided halves at the y/z/x(right) edges
    assert np.allclose(gy[:, 0, :], b / 2) and np.allclose(gy[:, -1, :], b / 2)
    assert np.allclose(gz[0], c / 2) and np.allclose(gz[-1], c / 2)
    assert np.allclose(gx[:, :, -1], a / 2)
    # x = 0 wraps into the previous row: left operand is f(w-1, y-1, z)
    assert np.allclose(gx[:, 1:, 0], (a * (2 - w) + b) / 2)
    # ... and at y = 0, z > 0 into the last row of the previous plane
    assert np.allclose(gx[1:, 0, 0], (a - a * (w - 1) - b * (h - 1) + c) / 2)
    # (0,0,0) is defined as the clamp
    assert np.isclose(gx[0, 0, 0], a / 2)


def test_oracle_gradient_matches_numpy_linear_index_form():
    rng = np.random.default_rng(2)
    f = rng.normal(size=(5, 6, 7)).astype(np.float32)
    for axis, got in enumerate(oracle.gradient(f)):
        np.testing.assert_allclose(got, npr.diff(f.astype(np.float64), axis), rtol=1e-6, atol=1e-7)


def test_oracle_diff_transpose_identity():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(5, 6, 7))
    u = rng.normal(size=(5, 6, 7))
    for axis in range(3):
        lhs = np.vdot(oracle.diff_f64(f, axis), u)
        rhs = np.vdot(f, oracle.diff_transpose_f64(u, axis))
        assert abs(lhs - rhs) < 1e-12 * max(1, abs(lhs))
        np.testing.assert_allclose(oracle.diff_f64(f, axis), npr.diff(f, axis), rtol=0, atol=1e-14)


# ---------------------------------------------------- valid blur (kernel.cu:1000-1121)
def test_oracle_conv_brute_force():
    rng = np.random.default_rng(4)
    g, P, m = oracle.taps(1.0)
    v = rng.uniform(0, 255, size=(9, 10, 11)).astype(np.float32)
    got = oracle.conv_separable(v, g)
    k3 = np.einsum("i,j,k->ijk", g, g, g).astype(np.float64)
    ol, oh, ow = v.shape[0] - P + 1, v.shape[1] - P + 1, v.shape[2] - P + 1
    ref = np.empty((ol, oh, ow))
    for z in range(ol):
        for y in range(oh):
            for x in range(ow):
                ref[z, y, x] = np.sum(v[z:z + P, y:y + P, x:x + P] * k3)
    np.testing.assert_allclose(got, ref, rtol=2e-6)


def test_oracle_conv_constant_and_delta():
    g, P, m = oracle.taps(3.0)
    c = 37.0
    got = oracle.conv_separable(np.full((23, 22, 24), c, np.float32), g)
    np.testing.assert_allclose(got, c * 0.998686813, rtol=2e-6)
    d = np.zeros((21, 21, 21), np.float32)
    d[10, 10, 10] = 1.0
    v = np.zeros((41, 41, 41), np.float32)
    v[20, 20, 20] = 1.0
    got = oracle.conv_separable(v, g)  # 21^3 output equals the flipped outer product
    ref = np.einsum("i,j,k->ijk", g, g, g)[::-1, ::-1, ::-1]
    np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-12)


def test_oracle_pad_border():
    v = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    p = oracle.pad_border(v, 2)
    assert p.shape == (6, 7, 8)
    assert np.array_equal(p[2:4, 2:5, 2:6], v)
    band = np.ones(p.shape, bool)
    band[2:4, 2:5, 2:6] = False
    assert np.all(p[band] == 30.0)


# --------------------------------------------------------- energies (kernel.cu:25-55)
def test_oracle_energies_pointwise_and_window():
    rng = np.random.default_rng(5)
    I = rng.uniform(0, 255, (6, 7, 8)).astype(np.float32)
    fo = rng.uniform(0, 255, I.shape).astype(np.float32)
    fi = rng.uniform(0, 255, I.shape).astype(np.float32)
    Eo, Ei = oracle.energies(fo, fi, I, 0)
    assert np.array_equal(Eo, (I - fo) * (I - fo)) and np.array_equal(Ei, (I - fi) * (I - fi))
    Eo, Ei = oracle.energies(fo, fi, I, 1)
    ro, ri = npr.energies(I.astype(np.float64), fo.astype(np.float64), fi.astype(np.float64), 1)
    np.testing.assert_allclose(Eo, ro, rtol=1e-5)
    np.testing.assert_allclose(Ei, ri, rtol=1e-5)


# --------------------------------------------------------------- update (kernel.cu:384)
def test_oracle_final_phi_closed_form():
    args = [np.array([v], np.float32) for v in (1.5, 0.3, 400.0, 100.0, -0.2, 0.7)]
    out = oracle.final_phi(*args, 1.0, 1.0, 0.1, 65.025, 1.0)
    phi, dh, Eo, Ei, k, L = (float(a[0]) for a in args)
    ref = phi - dh * 1.0 * (Eo - Ei) * 0.1 + 65.025 * 0.1 * dh * k + 1.0 * 0.1 * (L - k)
    assert abs(out[0] - ref) < 1e-5
    # lambda1 != lambda2 weights the two energies separately
    out2 = oracle.final_phi(*args, 2.0, 0.5, 0.1, 65.025, 1.0)
    ref2 = phi - dh * (2.0 * Eo - 0.5 * Ei) * 0.1 + 65.025 * 0.1 * dh * k + 0.1 * (L - k)
    assert abs(out2[0] - ref2) < 1e-4


# ------------------------------------------------------------ whole step (main.cpp:234-300)
@pytest.mark.parametrize("sigma,eps,shape", [(1.0, 0.2, (24, 26, 28)), (3.0, 1.0, (30, 31, 33)),
                                             (1.0, 1.0, (22, 23, 25))])
def test_oracle_step_matches_numpy_fp64(sigma, eps, shape):
    img, phi = small_scene(shape, seed=int(sigma * 10))
    p = oracle.make_params(sigma=sigma, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=eps)
    got = oracle.step_debug(phi, img, p)
    g32, _, _ = oracle.taps(sigma)
    ref = npr.step(phi, img, sigma=sigma, eps=eps, g=g32.astype(np.float64), debug=True)
    # the fitting force moves phi by hundreds near the front, so FP32 rounding
    # scales with the largest |phi'|, not with |phi0|
    scale = np.abs(ref["phi"]).max()
    assert np.abs(got["phi"] - ref["phi"]).max() <= 2e-6 * scale
    np.testing.assert_allclose(got["fo"], ref["fo"], rtol=2e-5)
    np.testing.assert_allclose(got["fi"], ref["fi"], rtol=2e-5)
    np.testing.assert_allclose(got["kappa"], ref["kappa"], atol=2e-5)
    np.testing.assert_allclose(got["L"], ref["L"], atol=2e-5)
    # the pad band has f == 1 exactly (30/30)
    _, P, m = oracle.taps(sigma)
    band = np.ones(shape, bool)
    band[m:-m, m:-m, m:-m] = False
    assert np.all(got["fo"][band] == 1.0) and np.all(got["fi"][band] == 1.0)


def test_oracle_step_sigma2_window():
    img, phi = small_scene((22, 23, 24), seed=7)
    p = oracle.make_params(sigma=1.0, epsilon=1.0, sigma2=1.0)
    got = oracle.step(phi, img, p)
    g32, _, _ = oracle.taps(1.0)
    ref = npr.step(phi, img, sigma=1.0, eps=1.0, sigma2=1.0, g=g32.astype(np.float64))
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


def test_oracle_step_f64_agrees_with_fp32():
    img, phi = small_scene((22, 23, 24), seed=8)
    p = oracle.make_params(sigma=1.0, epsilon=1.0)
    a = oracle.step(phi, img, p)
    b = oracle.step_f64(phi, img, p)
    assert np.abs(a - b).max() <= 2e-6 * np.abs(b).max()


def test_oracle_constant_image_zero_force():
    _, phi = small_scene((24, 24, 24), seed=9)
    img = np.full(phi.shape, 77, np.uint8)
    p = oracle.make_params(sigma=1.0, epsilon=1.0)
    d = oracle.step_debug(phi, img, p)
    assert np.abs(d["F"]).max() < 1e-6 * 77 * 77


def test_oracle_sphere_curvature():
    n = 41
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    r = np.sqrt((x - 20.0) ** 2 + (y - 20.0) ** 2 + (z - 20.0) ** 2)
    phi = (r - 10.0).astype(np.float32)
    img = np.full(phi.shape, 50, np.uint8)
    d = oracle.step_debug(phi, img, oracle.make_params(sigma=1.0))
    shell = (r > 8) & (r < 14)
    np.testing.assert_allclose(d["kappa"][shell], 2 / r[shell], rtol=0.08)
    np.testing.assert_allclose(d["L"][shell], 2 / r[shell], rtol=0.08)


def test_oracle_plane_fixed_point_interior():
    l, h, w = 20, 21, 22
    z, y, x = np.meshgrid(np.arange(l), np.arange(h), np.arange(w), indexing="ij")
    phi = (0.3 * x + 0.5 * y - 0.4 * z - 2).astype(np.float32)
    img = np.full(phi.shape, 90, np.uint8)
    out = oracle.step(phi, img, oracle.make_params(sigma=1.0))
    inner = (slice(3, -3), slice(3, -3), slice(3, -3))
    np.testing.assert_allclose(out[inner], phi[inner], atol=2e-5)


# ----------------------------------------------------------------- adjoint (FP64)
def _fd_setup(sigma2=0.0):
    img, phi = small_scene((16, 17, 18), seed=11, tubes=2)
    p = oracle.make_params(sigma=1.0, lambda1=1.0, lambda2=0.8, nu=6.5, mu=0.7, dt=0.1,
                           epsilon=1.0, sigma2=sigma2)
    rng = np.random.default_rng(12)
    c = rng.normal(size=phi.shape)
    return img.astype(np.float64), phi.astype(np.float64) + 0.01 * rng.normal(size=phi.shape), p, c


@pytest.mark.parametrize("sigma2", [0.0, 1.0])
def test_oracle_adjoint_matches_central_differences(sigma2):
    img, phi, p, c = _fd_setup(sigma2)
    gin, gp = oracle.step_adjoint_f64(phi, img, p, c)
    rng = np.random.default_rng(13)
    flat = rng.choice(phi.size, 24, replace=False)
    h = 1e-4
    for f in flat:
        e = np.zeros(phi.size)
        e[f] = h
        e = e.reshape(phi.shape)
        fd = (np.vdot(c, oracle.step_f64(phi + e, img, p)) -
              np.vdot(c, oracle.step_f64(phi - e, img, p))) / (2 * h)
        assert abs(fd - gin.ravel()[f]) <= 1e-5 * max(1.0, abs(fd)), (f, fd, gin.ravel()[f])
    for k, name in enumerate(("lambda1", "lambda2", "nu", "mu")):
        hp = 1e-4
        pp = oracle.make_params(**{**_pdict(p), name: getattr(p, name) + hp})
        pm = oracle.make_params(**{**_pdict(p), name: getattr(p, name) - hp})
        # float32 parameters: use the values actually stored
        dp = getattr(pp, name) - getattr(pm, name)
        fd = (np.vdot(c, oracle.step_f64(phi, img, pp)) - np.vdot(c, oracle.step_f64(phi, img, pm))) / dp
        assert abs(fd - gp[k]) <= 1e-5 * max(1.0, abs(fd)), (name, fd, gp[k])


def _pdict(p):
    return dict(sigma=p.sigma, lambda1=p.lambda1, lambda2=p.lambda2, nu=p.nu, mu=p.mu, dt=p.dt,
                epsilon=p.epsilon, sigma2=p.sigma2)


def test_oracle_atan_matches_numpy():
    """The oracle's float atanf comes from a double arctangent built from IEEE
    basic operations; pinned against numpy's arctan in double (relative error)
    and in float (the rounded results agree)."""
    rng = np.random.default_rng(11)
    x = np.concatenate([
        np.linspace(-4, 4, 200001),
        rng.uniform(-1, 1, 200000),
        np.sign(rng.normal(size=200000)) * 10.0 ** rng.uniform(-12, 12, 200000),
        [0.0, -0.0, 0.66, -0.66, 2.41421356237309504880, np.inf, -np.inf, 1e-300, 1e300],
    ])
    got = oracle.atan_f64(x)
    ref = np.arctan(x)
    nz = ref != 0
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    assert rel.max() < 4e-16, rel.max()
    assert np.all(got[~nz] == 0)
    xf = x[np.abs(x) < 1e30].astype(np.float32)
    gotf = oracle.atan_f64(xf.astype(np.float64)).astype(np.float32)
    reff = np.arctan(xf.astype(np.float64)).astype(np.float32)
    assert np.mean(gotf != reff) < 1e-5
