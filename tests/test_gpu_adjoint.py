"""GPU adjoint of one RSF step and the unrolled autograd wrapper, checked
against the FP64 CPU adjoint (oracle) and FP64 central finite differences."""
import numpy as np
import pytest

import oracle
import rsf_amd
from gpu_util import oracle_params, scene, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch  # noqa: F401

    c = rsf_amd.Context(0)
    yield c
    c.close()


def _params(sigma, **kw):
    # FAST unless a case asks for LITERAL: the mode the unrolled training path runs
    base = dict(sigma=sigma, lambda1=1.0, lambda2=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0,
                mode=rsf_amd.MODE_FAST)
    base.update(kw)
    return rsf_amd.make_params(**base)


@pytest.mark.parametrize("sigma,shape,kw", [
    (1.0, (24, 26, 28), {}),
    (3.0, (34, 36, 38), {}),
    (3.0, (30, 33, 35), dict(lambda1=1.3, lambda2=0.7, nu=6.5, mu=0.6)),
    (3.0, (30, 33, 35), dict(mode=rsf_amd.MODE_LITERAL, epsilon=0.2)),
])
def test_adjoint_one_step_matches_fp64(ctx, sigma, shape, kw):
    import torch

    img, phi0, _ = scene(shape, tubes=5, seed=21)
    p = _params(sigma, **kw)
    u = np.random.default_rng(5).normal(size=shape).astype(np.float32)
    l, h, w = shape
    d_gin = torch.empty(shape, dtype=torch.float32, device="cuda")
    d_gp = torch.zeros(4, dtype=torch.float32, device="cuda")
    ctx.adjoint([(0, w, h, l)], to_dev(img), to_dev(phi0), to_dev(u), d_gin, d_gp, p)
    ref, ref_gp = oracle.step_adjoint_f64(phi0, img, oracle_params(p), u)
    got = d_gin.cpu().numpy()
    rel = float(np.abs(got - ref).max() / np.abs(ref).max())
    assert rel <= 1e-3, rel
    gp = d_gp.cpu().numpy()
    assert np.all(np.abs(gp - ref_gp) <= 1e-3 * np.abs(ref_gp) + 1e-6 * np.abs(ref_gp).max()), (gp, ref_gp)


def test_adjoint_blocks_equal_single(ctx):
    import torch

    shapes = [(26, 28, 30), (30, 27, 25), (24, 24, 24)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    for k, ((off, w, h, l), shp) in enumerate(zip(blocks, shapes)):
        bi, bp, _ = scene(shp, tubes=3, seed=40 + k)
        img[off:off + bi.size] = bi.ravel()
        phi[off:off + bp.size] = bp.ravel()
    u = np.random.default_rng(6).normal(size=total).astype(np.float32)
    p = _params(1.0)
    gin = torch.empty(total, device="cuda")
    gp = torch.zeros(4, device="cuda")
    ctx.adjoint(blocks, to_dev(img), to_dev(phi), to_dev(u), gin, gp, p)
    gp_sum = np.zeros(4)
    for (off, w, h, l) in blocks:
        n = w * h * l
        g1 = torch.empty(n, device="cuda")
        p1 = torch.zeros(4, device="cuda")
        ctx.adjoint([(0, w, h, l)], to_dev(img[off:off + n]), to_dev(phi[off:off + n]), to_dev(u[off:off + n]),
                    g1, p1, p)
        assert torch.equal(g1, gin[off:off + n])
        gp_sum += p1.cpu().numpy()
    np.testing.assert_allclose(gp.cpu().numpy(), gp_sum, rtol=1e-5)


@pytest.mark.parametrize("u8", [True, False])
def test_adjoint_uniform_batch_equals_single(ctx, u8):
    """Blocks of one shape on 4-voxel boundaries run every adjoint kernel once for
    the whole batch (block index in the grid): the same bits as each block alone."""
    import torch

    shape = (36, 34, 40)
    blocks, total = rsf_amd.pack_blocks([shape] * 3)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    for k, (off, w, h, l) in enumerate(blocks):
        bi, bp, _ = scene(shape, tubes=4, seed=60 + k)
        img[off:off + bi.size] = bi.ravel()
        phi[off:off + bp.size] = bp.ravel()
    u = np.random.default_rng(9).normal(size=total).astype(np.float32)
    p = _params(3.0, nu=30.0)
    conv = (lambda a: a) if u8 else (lambda a: a.astype(np.float32))
    gin = torch.empty(total, device="cuda")
    gp = torch.zeros(4, device="cuda")
    ctx.adjoint(blocks, to_dev(conv(img)), to_dev(phi), to_dev(u), gin, gp, p)
    gp_sum = np.zeros(4)
    for (off, w, h, l) in blocks:
        n = w * h * l
        g1 = torch.empty(n, device="cuda")
        p1 = torch.zeros(4, device="cuda")
        ctx.adjoint([(0, w, h, l)], to_dev(conv(img[off:off + n])), to_dev(phi[off:off + n]), to_dev(u[off:off + n]),
                    g1, p1, p)
        assert torch.equal(g1, gin[off:off + n])
        gp_sum += p1.cpu().numpy()
    np.testing.assert_allclose(gp.cpu().numpy(), gp_sum, rtol=1e-5)


@pytest.mark.parametrize("shape,u8", [((36, 40, 44), True), ((32, 30, 48), False), ((40, 36, 128), True)])
def test_adjoint_quad_kernels_equal_scalar(ctx, shape, u8):
    """Rows a multiple of 4 long on 16-byte aligned buffers run the stencil and
    finishing kernels four voxels per thread; the same block one voxel further on
    runs them a voxel per thread.  dphi must be the same bits, the parameter
    gradients the same up to the order of their partial sums."""
    import torch

    img, phi0, _ = scene(shape, tubes=5, seed=23)
    n = img.size
    u = np.random.default_rng(8).normal(size=n).astype(np.float32)
    p = _params(3.0, lambda1=1.2, lambda2=0.9)
    l, h, w = shape
    out = []
    for shift in (0, 1):
        def pad(a):
            b = np.zeros(n + shift, a.dtype)
            b[shift:] = a.ravel()
            return b
        d_img = to_dev(pad(img) if u8 else pad(img.astype(np.float32)))
        gin = torch.zeros(n + shift, device="cuda")
        gp = torch.zeros(4, device="cuda")
        ctx.adjoint([(shift, w, h, l)], d_img, to_dev(pad(phi0)), to_dev(pad(u)), gin, gp, p)
        out.append((gin[shift:].cpu().numpy(), gp.cpu().numpy()))
    (g0, p0), (g1, p1) = out
    assert np.array_equal(g0.view(np.uint32), g1.view(np.uint32))
    np.testing.assert_allclose(p0, p1, rtol=1e-5)


def _fwd64(phi, img, p, K):
    st = [phi.astype(np.float64)]
    for _ in range(K):
        st.append(oracle.step_f64(st[-1], img, p))
    return st


def _loss_block64(phiK, mask, eps):
    z = -phiK / eps
    return float(np.sum(np.logaddexp(0.0, z) - mask * z))  # BCE with logits, summed


def test_unrolled_gradient_matches_fp64_adjoint_and_central_differences(ctx):
    """K = 4 unrolled steps, checkpoints every 2, BCE of sigmoid(-phi/eps).

    Reference gradient: the FP64 CPU adjoint chained over the FP64 forward
    states.  Finite differences (FP64 forward, h = 1e-3) are compared at random
    voxels where they are converged, i.e. where FD(h) and FD(h/10) agree: next
    to voxels whose central-difference gradient is exactly zero, the normal
    g/(|g|+1e-6) has slope ~1e6 and a step of 1e-3 saturates it."""
    import torch

    from rsf_amd.autograd import bce_inside_loss, rsf_unrolled

    shapes = [(22, 24, 26), (24, 22, 25)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    imgs, phis, masks = [], [], []
    for k, shp in enumerate(shapes):
        bi, bp, bm = scene(shp, tubes=4, seed=60 + k)
        imgs.append(bi)
        phis.append(bp)
        masks.append(bm.astype(np.float64))
    img = np.concatenate([a.ravel() for a in imgs])
    phi = np.concatenate([a.ravel() for a in phis])
    mask = np.concatenate([a.ravel() for a in masks]).astype(np.float32)
    K, C, eps = 4, 2, 1.0
    theta0 = [1.0, 1.0, 6.5, 0.5]
    d_phi = to_dev(phi).requires_grad_(True)
    theta = torch.tensor(theta0, dtype=torch.float32, device="cuda", requires_grad=True)
    out = rsf_unrolled(d_phi, theta, to_dev(img), blocks, K, C, sigma=1.0, epsilon=eps, context=ctx)
    loss = bce_inside_loss(out, to_dev(mask), eps)
    loss.backward()
    g = d_phi.grad.cpu().numpy()
    gth = theta.grad.cpu().numpy()

    op = oracle.make_params(sigma=1.0, lambda1=theta0[0], lambda2=theta0[1], nu=theta0[2], mu=theta0[3],
                            dt=0.1, epsilon=eps)
    states = [_fwd64(phis[b], imgs[b], op, K) for b in range(2)]
    L0 = sum(_loss_block64(states[b][-1], masks[b], eps) for b in range(2)) / total
    assert abs(L0 - float(loss.detach())) <= 1e-4 * abs(L0)
    # FP64 adjoint chain
    ref = np.empty(total)
    ref_th = np.zeros(4)
    for b, (off, w, h, l) in enumerate(blocks):
        zK = -states[b][-1] / eps
        u = (1.0 / (1.0 + np.exp(-zK)) - masks[b]) * (-1.0 / eps) / total
        for t in reversed(range(K)):
            u, gp = oracle.step_adjoint_f64(states[b][t], imgs[b], op, u)
            ref_th += gp
        ref[off:off + w * h * l] = u.ravel()
    err = np.abs(g - ref)
    assert float(np.linalg.norm(g - ref) / np.linalg.norm(ref)) <= 1e-3
    assert float(err.max()) <= 1e-3 * float(np.abs(ref).max()), float(err.max() / np.abs(ref).max())
    np.testing.assert_allclose(gth, ref_th, rtol=1e-3)

    def loss_with(b, v, dv):
        pb = phis[b].astype(np.float64).ravel().copy()
        pb[v] += dv
        st = _fwd64(pb.reshape(shapes[b]), imgs[b], op, K)
        other = 1 - b
        return (_loss_block64(st[-1], masks[b], eps) + _loss_block64(states[other][-1], masks[other], eps)) / total

    rng = np.random.default_rng(7)
    checked = 0
    for v in rng.choice(total, 24, replace=False):
        b = 0 if v < blocks[1][0] else 1
        lv = int(v - blocks[b][0])
        fd = [(loss_with(b, lv, hh) - loss_with(b, lv, -hh)) / (2 * hh) for hh in (1e-3, 1e-4)]
        scale = max(abs(fd[1]), 1e-3 * float(np.abs(ref).max()))
        if abs(fd[0] - fd[1]) > 1e-3 * scale:
            continue  # FD not converged at h = 1e-3 here
        checked += 1
        assert abs(g[v] - fd[0]) <= 1e-3 * scale, (int(v), float(g[v]), fd[0])
        assert abs(ref[v] - fd[0]) <= 1e-4 * scale, (int(v), float(ref[v]), fd[0])
    assert checked >= 18, checked
    for k, name in enumerate(("lambda1", "lambda2", "nu", "mu")):
        hk = 1e-6 * max(1.0, abs(theta0[k]))
        vals = []
        for sgn in (1, -1):
            t = list(theta0)
            t[k] += sgn * hk
            pk = oracle.make_params(sigma=1.0, lambda1=t[0], lambda2=t[1], nu=t[2], mu=t[3], dt=0.1, epsilon=eps)
            # float32 parameter storage: difference of the stored values
            vals.append((sum(_loss_block64(_fwd64(phis[b], imgs[b], pk, K)[-1], masks[b], eps)
                             for b in range(2)) / total, getattr(pk, name)))
        fd = (vals[0][0] - vals[1][0]) / (vals[0][1] - vals[1][1])
        # lambda moves phi by ~dt*delta*F ~ 1e3 per unit: even a 1e-6 step (the smallest the
        # float32 parameter holds) is not fully linear at the degenerate normals, so 1e-3 here
        assert abs(ref_th[k] - fd) <= 1e-3 * max(abs(fd), 1e-3 * float(np.abs(ref_th).max())), (name, ref_th[k], fd)


def test_held_image_gives_identical_results(ctx):
    """rsf_context_hold_image reuses the image's preparation across calls: the
    steps and the adjoint give the same bits as without it, and after a release
    a changed image is prepared afresh."""
    import torch

    shapes = [(30, 32, 34), (28, 30, 36)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    for k, ((off, w, h, l), shp) in enumerate(zip(blocks, shapes)):
        bi, bp, _ = scene(shp, tubes=4, seed=70 + k)
        img[off:off + bi.size] = bi.ravel()
        phi[off:off + bp.size] = bp.ravel()
    p = _params(3.0)
    d_img = to_dev(img)
    u = to_dev(np.random.default_rng(8).normal(size=total).astype(np.float32))

    def run(hold):
        if hold:
            ctx.hold_image(d_img)
        try:
            a, b = to_dev(phi), torch.empty(total, device="cuda")
            for _ in range(3):
                ctx.step_blocks(blocks, d_img, a, b, p)
                a, b = b, a
            gin = torch.empty(total, device="cuda")
            gp = torch.zeros(4, device="cuda")
            ctx.adjoint(blocks, d_img, a, u, gin, gp, p)
            torch.cuda.synchronize()
            return a.cpu().numpy(), gin.cpu().numpy(), gp.cpu().numpy()
        finally:
            if hold:
                ctx.release_image()

    ref = run(False)
    got = run(True)
    for r, g in zip(ref, got):
        assert np.array_equal(r, g)
    # released: an edited image is prepared again
    ctx.hold_image(d_img)
    run(False)
    ctx.release_image()
    d_img[:100] = 255
    img2 = d_img.cpu().numpy()
    fresh = rsf_amd.Context(0)
    a1, b1 = to_dev(phi), torch.empty(total, device="cuda")
    a2, b2 = to_dev(phi), torch.empty(total, device="cuda")
    ctx.step_blocks(blocks, d_img, a1, b1, p)
    fresh.step_blocks(blocks, to_dev(img2), a2, b2, p)
    torch.cuda.synchronize()
    fresh.close()
    assert np.array_equal(b1.cpu().numpy(), b2.cpu().numpy())


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_unrolled_forward_equals_steps(ctx, mode):
    """The unrolled forward runs each checkpoint segment as one graph-replayed
    batched run; the backward recomputes the segments step by step.  Both give
    the same bits: K = 11 with checkpoints every 4 (a short last segment)."""
    import torch

    from rsf_amd.autograd import rsf_unrolled

    shapes = [(30, 32, 34), (28, 30, 36)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    for k, ((off, w, h, l), shp) in enumerate(zip(blocks, shapes)):
        bi, bp, _ = scene(shp, tubes=4, seed=80 + k)
        img[off:off + bi.size] = bi.ravel()
        phi[off:off + bp.size] = bp.ravel()
    K, C = 11, 4
    d_img = to_dev(img)
    theta = torch.tensor([1.0, 1.0, 65.025, 1.0], device="cuda")
    out = rsf_unrolled(to_dev(phi), theta, d_img, blocks, K, C, sigma=3.0, epsilon=1.0, mode=mode, context=ctx)
    p = _params(3.0, mode=mode)
    a, b = to_dev(phi), torch.empty(total, device="cuda")
    for _ in range(K):
        ctx.step_blocks(blocks, d_img, a, b, p)
        a, b = b, a
    torch.cuda.synchronize()
    assert np.array_equal(out.detach().cpu().numpy(), a.cpu().numpy())
