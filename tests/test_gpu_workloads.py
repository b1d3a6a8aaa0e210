"""GPU parity at the benchmark's batched workload: 64 uneven blocks with
sides from {96, 160, 192, 256} (350 Mvoxel), the generator and block set the
benchmark times (bench.batched_blocks), run with the kernels rsf_autotune picks
for that block set -- the replacement of one rsf3d process per block
(README.md:56, CUDA_RSF/main.cpp:222-300 per block)."""
import numpy as np
import pytest

import oracle
import rsf_amd
from gpu_util import one_step_gate, oracle_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def blockset():
    import bench

    return bench.batched_blocks(rsf_amd, 64)


def _pick(shapes):
    """Largest, smallest and two edge-like (thin along one axis) blocks."""
    vol = [a * b * c for a, b, c in shapes]
    idx = [int(np.argmax(vol)), int(np.argmin(vol))]
    for k in range(len(shapes)):
        l, h, w = shapes[k]
        thin = min(shapes[k]) == 96 and max(shapes[k]) == 256
        if thin and k not in idx:
            idx.append(k)
        if len(idx) == 4:
            break
    return idx


@pytest.mark.parametrize("mode", [rsf_amd.MODE_LITERAL, rsf_amd.MODE_FAST])
def test_batched_workload_autotuned_equals_sequential_and_oracle(blockset, mode):
    """Two iterations of the 64-block set in one persistent launch per kernel with
    the autotuned tile shapes equal the blocks run one after another through the
    single-volume engine, bit for bit; the largest, the smallest and two thin
    blocks equal the oracle's run on that block alone (LITERAL bit for bit, FAST
    within the one-step gate)."""
    import torch

    shapes, blocks, img, phi = blockset
    print(f"batched workload mode={mode}", flush=True)
    ctx = rsf_amd.Context(0)
    try:
        p = rsf_amd.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0, mode=mode,
                                iterations=2)
        choice = ctx.autotune(blocks, rsf_amd.DTYPE_U8, p)
        print(f"autotune: {choice}", flush=True)
        d_img = torch.from_numpy(img).cuda()
        a = torch.from_numpy(phi).cuda()
        b = a.clone()
        ctx.run_batched(blocks, d_img, a, p)
        ctx.run_sequential(blocks, d_img, b, p)
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        A = a.cpu().numpy()
    finally:
        ctx.close()
    idx = _pick(shapes)
    assert len(idx) == 4
    for k in idx:
        off, w, h, l = blocks[k]
        n = w * h * l
        shp = shapes[k]
        ref = oracle.run(phi[off:off + n].reshape(shp), img[off:off + n].reshape(shp).astype(np.float32),
                         oracle_params(p), 2)
        got = A[off:off + n].reshape(shp)
        if mode == rsf_amd.MODE_LITERAL:
            assert np.array_equal(got, ref), (k, shp, float(np.abs(got - ref).max()))
        else:
            err, tol = one_step_gate(got, ref)
            assert err <= tol, (k, shp, err, tol)


def _config4_batch(nvol=8, n=128):
    """The differentiated workload: nvol tube phantoms of n^3 (40 tubes, seeds
    4000+b), phi0 the signed distance of the thresholded sigma-1 blur, and the
    generator's 0/1 masks."""
    shapes = [(n, n, n)] * nvol
    blocks, total = rsf_amd.pack_blocks(shapes)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    mask = np.empty(total, np.float32)
    for b, (off, w, h, l) in enumerate(blocks):
        bi, bm = rsf_amd.synth_tubes((n, n, n), tubes=40, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=4000 + b)
        img[off:off + n ** 3] = bi.ravel()
        mask[off:off + n ** 3] = bm.ravel()
        phi[off:off + n ** 3] = rsf_amd.initial_phi(bi, 1.0, 110.0).ravel()
    return blocks, img, phi, mask


def _fp64_adjoint_along(states, img, mask64, op, total, eps):
    """dL/dphi0 and dL/dtheta by the FP64 adjoint of each step, evaluated at the
    given forward states, for L = sum of BCE(sigmoid(-phi_K/eps), mask) / total."""
    zK = -states[-1].astype(np.float64) / eps
    u = (1.0 / (1.0 + np.exp(-zK)) - mask64) * (-1.0 / eps) / total
    th = np.zeros(4)
    for t in reversed(range(len(states) - 1)):
        u, gp = oracle.step_adjoint_f64(states[t], img, op, u)
        th += gp
    return u, th


def test_differentiated_workload_gradient_matches_fp64_adjoint():
    """8 x 128^3, K = 32 unrolled steps with checkpoints every 8, BCE of
    sigmoid(-phi/eps) against the masks, LITERAL mode.

    At nu*dt = 6.5 the explicit scheme amplifies rounding: an FP64 forward and
    the reference's FP32 forward part ways (max|dphi|/max|phi| ~ 0.2 at step 16),
    and FP64 adjoints along the two trajectories differ by ~4 % in the phi0
    gradient and more in the parameter gradients.  The device forward in LITERAL
    mode is the oracle's FP32 forward bit for bit (checked here for phi_K), so
    the reference gradient is the FP64 adjoint of every step chained along those
    same FP32 states: the comparison isolates the device adjoint's own FP32
    arithmetic.  (The FP64 adjoint chain is checked against FP64 central
    differences at this size in tests/test_config4_fd.py.)"""
    import torch

    from rsf_amd.autograd import bce_inside_loss, rsf_unrolled

    K, C, eps = 32, 8, 1.0
    blocks, img, phi, mask = _config4_batch()
    total = img.size
    n = blocks[0][1] * blocks[0][2] * blocks[0][3]
    shape = (blocks[0][3], blocks[0][2], blocks[0][1])
    theta0 = [1.0, 1.0, 65.025, 1.0]
    ctx = rsf_amd.Context(0)
    try:
        dev = torch.device("cuda:0")
        d_img, d_mask = torch.from_numpy(img).to(dev), torch.from_numpy(mask).to(dev)
        d_phi = torch.from_numpy(phi).to(dev).requires_grad_(True)
        theta = torch.tensor(theta0, device=dev, requires_grad=True)
        out = rsf_unrolled(d_phi, theta, d_img, blocks, K, C, sigma=3.0, epsilon=eps, mode=rsf_amd.MODE_LITERAL,
                           context=ctx)
        loss = bce_inside_loss(out, d_mask, eps)
        loss.backward()
        g = d_phi.grad.cpu().numpy()
        phiK = out.detach().cpu().numpy()
        # theta's gradient of volumes 0 and 1 alone (their loss normalised over the pair)
        d2 = torch.from_numpy(phi[:2 * n]).to(dev).requires_grad_(True)
        th2 = torch.tensor(theta0, device=dev, requires_grad=True)
        out2 = rsf_unrolled(d2, th2, d_img[:2 * n], blocks[:2], K, C, sigma=3.0, epsilon=eps,
                            mode=rsf_amd.MODE_LITERAL, context=ctx)
        bce_inside_loss(out2, d_mask[:2 * n], eps).backward()
        gth2 = th2.grad.cpu().numpy()
    finally:
        ctx.close()
    op = oracle.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=eps)
    ref_th = np.zeros(4)
    errs = []
    for b in range(2):
        sl = slice(b * n, (b + 1) * n)
        I32 = img[sl].reshape(shape).astype(np.float32)
        states = [phi[sl].reshape(shape)]
        for _ in range(K):
            states.append(oracle.step(states[-1], I32, op))
        assert np.array_equal(states[-1], phiK[sl].reshape(shape)), b   # same FP32 trajectory
        m64 = mask[sl].reshape(shape).astype(np.float64)
        ref, _ = _fp64_adjoint_along(states, img[sl].reshape(shape), m64, op, total, eps)
        _, th = _fp64_adjoint_along(states, img[sl].reshape(shape), m64, op, 2 * n, eps)
        ref_th += th
        gb = g[sl].reshape(shape)
        errs.append((float(np.linalg.norm(gb - ref) / np.linalg.norm(ref)),
                     float(np.abs(gb - ref).max() / np.abs(ref).max())))
    th_rel = np.abs(gth2 - ref_th) / np.abs(ref_th)
    print(f"config4: phi0 gradient (rel l2, rel max) per volume {errs}; theta rel {th_rel}", flush=True)
    for e2, emax in errs:
        assert e2 <= 1e-3 and emax <= 1e-3, (errs, th_rel.tolist())
    assert np.all(th_rel <= 1e-3), (errs, th_rel.tolist(), gth2.tolist(), ref_th.tolist())
