"""GPU parity of the fused engine (sweeps, graph-captured loop, batched
blocks) against the oracle, with the parity gates: after one step
max|dphi| <= 1e-4 * max|phi|; after a run the sign masks differ on <= 1e-4
of the voxels."""
import numpy as np
import pytest

import oracle
import rsf_amd
from gpu_util import mask_disagreement, one_step_gate, oracle_params, scene, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch  # noqa: F401  (binds librsf to torch's HIP runtime)

    c = rsf_amd.Context(0)
    yield c
    c.close()


CASES = [
    # sigma, eps, mode, image dtype, shape (l,h,w)
    (3.0, 1.0, rsf_amd.MODE_FAST, np.uint8, (48, 50, 53)),
    (3.0, 0.2, rsf_amd.MODE_FAST, np.uint8, (45, 64, 70)),
    (3.0, 1.0, rsf_amd.MODE_LITERAL, np.uint8, (48, 50, 53)),
    (3.0, 0.2, rsf_amd.MODE_LITERAL, np.float32, (41, 47, 66)),
    (1.0, 1.0, rsf_amd.MODE_FAST, np.float32, (30, 33, 37)),
    (1.0, 0.2, rsf_amd.MODE_LITERAL, np.uint8, (30, 33, 37)),
    (5.0, 1.0, rsf_amd.MODE_FAST, np.uint8, (40, 41, 43)),
    (5.0, 1.0, rsf_amd.MODE_LITERAL, np.uint8, (40, 41, 43)),
    (1.3, 1.0, rsf_amd.MODE_FAST, np.uint8, (24, 26, 28)),   # P = 9
    (2.5, 1.0, rsf_amd.MODE_LITERAL, np.uint8, (36, 30, 52)),  # P = 17, u = 12 (effective sigma sqrt 6)
    (2.5, 0.2, rsf_amd.MODE_FAST, np.float32, (33, 35, 38)),
    (4.8, 1.0, rsf_amd.MODE_LITERAL, np.uint8, (40, 41, 43)),  # P = 33
    (5.3, 1.0, rsf_amd.MODE_FAST, np.uint8, (40, 41, 43)),     # P = 37: unfused fallback
    # w a multiple of 4: the vector interior of sweep 2
    (3.0, 1.0, rsf_amd.MODE_FAST, np.uint8, (40, 44, 48)),
    (3.0, 0.2, rsf_amd.MODE_LITERAL, np.uint8, (36, 30, 52)),
    (1.0, 1.0, rsf_amd.MODE_FAST, np.float32, (30, 33, 36)),
    (5.0, 1.0, rsf_amd.MODE_FAST, np.uint8, (40, 41, 44)),
    (3.0, 1.0, rsf_amd.MODE_FAST, np.uint8, (23, 22, 24)),  # one row of quads per tile
    # P = 21 blocks a few columns wider than the blur: the narrow-block strip kernels
    (3.0, 1.0, rsf_amd.MODE_LITERAL, np.uint8, (23, 22, 24)),
    (3.0, 1.0, rsf_amd.MODE_LITERAL, np.float32, (26, 23, 25)),
    (3.0, 1.0, rsf_amd.MODE_FAST, np.uint8, (26, 23, 25)),
]


@pytest.mark.parametrize("sigma,eps,mode,dtype,shape", CASES)
def test_fused_one_step_matches_oracle(ctx, sigma, eps, mode, dtype, shape):
    import torch

    img, phi0, _ = scene(shape, tubes=6, seed=int(sigma * 10) + mode)
    p = rsf_amd.make_params(sigma=sigma, epsilon=eps, mode=mode, iterations=1)
    d_img = to_dev(img.astype(dtype))
    d_phi = to_dev(phi0)
    d_out = torch.empty_like(d_phi)
    ctx.step(d_img, d_phi, d_out, p)
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    ref = oracle.step(phi0, img.astype(np.float32), oracle_params(p))
    err, tol = one_step_gate(got, ref)
    assert err <= tol, (err, tol)
    if mode == rsf_amd.MODE_LITERAL:
        # the literal path performs the oracle's operations in the oracle's order
        assert np.array_equal(got, ref), err
    else:
        # fast mode: the same sums regrouped (H = 1/2 + G), FP32 rounding apart
        assert err <= 2e-5 * float(np.abs(ref).max()), err


def test_fused_lambda1_ne_lambda2_and_sigma2(ctx):
    import torch

    img, phi0, _ = scene((36, 38, 40), tubes=5, seed=3)
    for kw in (dict(lambda1=1.3, lambda2=0.7), dict(lambda1=1.3, lambda2=0.7, mode=rsf_amd.MODE_FAST),
               dict(sigma2=1.0), dict(sigma2=1.0, mode=rsf_amd.MODE_LITERAL)):
        p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=1, **kw)
        d_phi, d_img = to_dev(phi0), to_dev(img)
        d_out = torch.empty_like(d_phi)
        ctx.step(d_img, d_phi, d_out, p)
        got = d_out.cpu().numpy()
        ref = oracle.step(phi0, img.astype(np.float32), oracle_params(p))
        err, tol = one_step_gate(got, ref)
        assert err <= tol, (kw, err, tol)
        if "lambda2" in kw and p.mode == rsf_amd.MODE_LITERAL:
            # separate lambdas take their own force form (kernel.cu:44-45), still bit for bit
            assert np.array_equal(got, ref), (kw, err)


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_run_matches_oracle_mask(ctx, mode):
    """Config-1 style run (128^3 is the bench's CPU case; 64^3 x 40 here)."""
    img, phi0, _ = scene((64, 64, 64), tubes=12, seed=1)
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, mode=mode, iterations=40)
    d_img, d_phi = to_dev(img), to_dev(phi0)
    ctx.run(d_img, d_phi, p)
    got = d_phi.cpu().numpy()
    ref = oracle.run(phi0, img.astype(np.float32), oracle_params(p), 40)
    assert mask_disagreement(got, ref) <= 1e-4
    # away from the front the two agree closely; single voxels at the front can
    # drift after many steps (the gate above is the acceptance criterion)
    diff = np.abs(got - ref)
    assert float(np.median(diff)) <= 1e-5 * float(np.abs(ref).max())


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_graph_run_equals_eager_steps(ctx, mode):
    import torch

    img, phi0, _ = scene((40, 42, 44), tubes=5, seed=4)
    p1 = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=1, mode=mode)
    d_img = to_dev(img)
    a, b = to_dev(phi0), torch.empty_like(to_dev(phi0))
    eager = {}
    for t in range(1, 71):
        ctx.step(d_img, a, b, p1)
        a, b = b, a
        if t in (5, 70):
            eager[t] = a.cpu().numpy()
    # T = 5: one 4-iteration graph + remainder; T = 70: 64-iteration graph + 6
    for T in (5, 70):
        pT = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=T, mode=mode)
        for _ in range(2):  # the second call replays the cached graph
            d_phi = to_dev(phi0)
            ctx.run(d_img, d_phi, pT)
            assert np.array_equal(d_phi.cpu().numpy(), eager[T])


def test_per_iteration_timing_host_path(ctx):
    img, phi0, _ = scene((40, 40, 40), tubes=4, seed=5)
    for T in (1, 4, 7):
        p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=T)
        out, ms = ctx.run_host(img, phi0, p, per_iteration=True)
        assert ms.shape == (T,) and np.all(ms > 0)
        ref = oracle.run(phi0, img.astype(np.float32), oracle_params(p), T)
        assert mask_disagreement(out, ref) <= 1e-4


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_batched_equals_sequential_blocks(ctx, mode):
    import torch

    rng = np.random.default_rng(6)
    shapes = [tuple(int(s) for s in rng.choice([24, 33, 40, 52], 3)) for _ in range(7)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    for k, ((off, w, h, l), shp) in enumerate(zip(blocks, shapes)):
        bi, bp, _ = scene(shp, tubes=3, seed=100 + k)
        img[off:off + bi.size] = bi.ravel()
        phi[off:off + bp.size] = bp.ravel()
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=6, mode=mode)
    d_img = to_dev(img)
    a, b = to_dev(phi), to_dev(phi)
    ctx.run_batched(blocks, d_img, a, p)
    ctx.run_sequential(blocks, d_img, b, p)
    torch.cuda.synchronize()
    A, B = a.cpu().numpy(), b.cpu().numpy()
    assert np.array_equal(A, B)
    # and each block equals the oracle on that block alone: bit for bit in
    # literal mode, within the mask gate in fast mode
    for (off, w, h, l), shp in zip(blocks, shapes):
        ref = oracle.run(phi[off:off + w * h * l].reshape(shp), img[off:off + w * h * l].reshape(shp)
                         .astype(np.float32), oracle_params(p), 6)
        got = A[off:off + w * h * l].reshape(shp)
        if mode == rsf_amd.MODE_LITERAL:
            assert np.array_equal(got, ref), (shp, float(np.abs(got - ref).max()))
        else:
            assert mask_disagreement(got, ref) <= 1e-4


def test_unfused_device_step_matches_oracle(ctx):
    import torch

    img, phi0, _ = scene((30, 31, 33), tubes=4, seed=7)
    p = rsf_amd.make_params(sigma=3.0, epsilon=0.2, iterations=1, mode=rsf_amd.MODE_LITERAL)
    out = torch.empty_like(to_dev(phi0))
    ctx.step_unfused(to_dev(img.astype(np.float32)), to_dev(phi0), out, p)
    ref = oracle.step(phi0, img.astype(np.float32), oracle_params(p))
    err, _ = one_step_gate(out.cpu().numpy(), ref)
    assert err <= 1e-6 * np.abs(ref).max()


def test_errors_are_reported(ctx):
    import torch

    img, phi0, _ = scene((24, 24, 24), tubes=2, seed=8)
    d_img, d_phi = to_dev(img), to_dev(phi0)
    out = torch.empty_like(d_phi)
    with pytest.raises(rsf_amd.RSFError) as e:
        ctx.step(d_img, d_phi, out, rsf_amd.make_params(sigma=2.0))   # floor(7 sigma) even
    assert e.value.status == rsf_amd.ERR_INVALID_ARGUMENT
    with pytest.raises(rsf_amd.RSFError):
        ctx.step(d_img, d_phi, d_phi, rsf_amd.make_params(sigma=3.0))  # aliasing
    small = to_dev(phi0[:10, :10, :10].copy())
    with pytest.raises(rsf_amd.RSFError):
        ctx.step(to_dev(img[:10, :10, :10].copy()), small, torch.empty_like(small), rsf_amd.make_params(sigma=3.0))


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_unaligned_buffers_match_aligned(ctx, mode):
    """phi buffers that do not start on a 16-byte boundary take the scalar
    interior kernel; the result matches the aligned (vector) run."""
    import torch

    shape = (36, 40, 44)
    img, phi0, _ = scene(shape, tubes=5, seed=17)
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, mode=mode, iterations=1)
    n = phi0.size
    d_img = to_dev(img)
    base_in = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
    base_out = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
    u_in, u_out = base_in[1:].view(shape), base_out[1:].view(shape)
    u_in.copy_(to_dev(phi0))
    ctx.step(d_img, u_in, u_out, p)
    a_in, a_out = to_dev(phi0), torch.empty(shape, dtype=torch.float32, device="cuda")
    ctx.step(d_img, a_in, a_out, p)
    got_u, got_a = u_out.cpu().numpy(), a_out.cpu().numpy()
    ref = oracle.step(phi0, img.astype(np.float32), oracle_params(p))
    for got in (got_u, got_a):
        err, tol = one_step_gate(got, ref)
        assert err <= tol, (err, tol)
    assert np.abs(got_u - got_a).max() <= 1e-6 * np.abs(ref).max()


def _registered(mode, P=21):
    """Every kernel variant the library can run (and rsf_autotune can pick) for
    this tap count and mode: sweep-1 names, sweep-2 names."""
    lit = mode == rsf_amd.MODE_LITERAL
    s1 = [n for n, taps, l in rsf_amd.variants(1) if taps == P and l == lit]
    s2 = [n for n, _, l in rsf_amd.variants(2) if l == lit]
    return s1, s2


@pytest.mark.parametrize("shape", [(40, 44, 48), (52, 97, 130), (37, 131, 236), (29, 57, 301)],
                         ids=["40x44x48", "52x97x130", "37x131x236", "29x57x301"])
@pytest.mark.parametrize("lambdas", [(1.0, 1.0), (1.3, 0.7)])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_kernel_variants_give_identical_results(ctx, dtype, mode, lambdas, shape, monkeypatch):
    """Every registered variant -- each sweep-1 kernel for P = 21 (fused, split
    xy/z) and each sweep-2 kernel (vector, scalar buffer, direct, LDS), the set
    rsf_autotune chooses from -- computes the same operations in the same order:
    a run gives bit-identical phi whichever variants and z chunks it takes, with
    equal lambdas and with separate ones (whose force forms differ).  The fast
    mode's matrix-core xy kernels sum the blur in another order: they are held
    to the fast-mode gate against the oracle (3 steps within 1e-4 max|phi|).  The second
    shape has full and partial tiles and rows that are not 16-byte aligned (w =
    130), the first aligned rows (the vector sweep-2 interior); the third four
    64-column strips, the last partial, and a row count (131) that is not a
    multiple of the strip kernel's 21-row groups, with the strip kernel as the
    literal default (wide enough for it); the fourth an odd width (rows that do
    not start on 16-byte boundaries: the scalar sweep-2 interior) under the strips
    and a row count below three ring groups.  In literal mode the default
    run is also the oracle's phi bit for bit."""
    import torch

    img, phi0, _ = scene(shape, tubes=6, seed=23)
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=3, mode=mode, lambda1=lambdas[0],
                            lambda2=lambdas[1])
    d_img = to_dev(img.astype(dtype))
    s1, s2 = _registered(mode)
    combos = [(None, None)]
    combos += [(f"{n}:{zc}", None) for n in s1 for zc in (8, 32)]
    combos += [(None, f"{n}:{zc}") for n in s2 for zc in (8, 64)]
    results = {}
    for c1, c2 in combos:
        for var, val in (("RSF_SWEEP1", c1), ("RSF_SWEEP2", c2)):
            if val is None:
                monkeypatch.delenv(var, raising=False)
            else:
                monkeypatch.setenv(var, val)
        d_phi = to_dev(phi0)
        ctx.run(d_img, d_phi, p)
        torch.cuda.synchronize()
        results[(c1, c2)] = d_phi.cpu().numpy()
    ref = results[(None, None)]
    want = oracle.run(phi0, img.astype(np.float32), oracle_params(p), 3)
    for k, v in results.items():
        if k[0] is not None and "mfma" in k[0]:
            # the matrix-core blurs sum in their own order (fast mode only): the
            # fast-mode gate against the oracle instead of bit equality
            assert mode == rsf_amd.MODE_FAST
            err, _ = one_step_gate(v, want)
            assert err <= 1e-4 * float(np.abs(want).max()), (k, err)
            continue
        assert np.array_equal(v, ref), (k, float(np.abs(v - ref).max()))
    if mode == rsf_amd.MODE_LITERAL:
        assert np.array_equal(ref, want), float(np.abs(ref - want).max())
    else:
        err, _ = one_step_gate(ref, want)
        assert err <= 1e-4 * float(np.abs(want).max()), err


@pytest.mark.parametrize("P", list(range(7, 37, 2)))
@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_every_odd_tap_count_one_step(ctx, mode, P):
    """sigma = (P + 1/2)/7 gives floor(7 sigma) = P (main.cpp:152-155): every odd
    tap count from 7 to 35 runs its fused sweep-1 kernel (P = 21 the split
    default).  LITERAL: one step equals the oracle bit for bit; FAST: within
    2e-5 max|phi|.  The batched launch of two uneven blocks equals the blocks
    run one by one."""
    import torch

    sigma = float(np.float32((P + 0.5) / 7.0))
    assert rsf_amd.taps(sigma)[1] == P
    shape = (40, 41, 44)
    img, phi0, _ = scene(shape, tubes=6, seed=200 + P)
    p = rsf_amd.make_params(sigma=sigma, epsilon=1.0, mode=mode, iterations=1)
    d_out = torch.empty(shape, dtype=torch.float32, device="cuda")
    ctx.step(to_dev(img), to_dev(phi0), d_out, p)
    torch.cuda.synchronize()
    got = d_out.cpu().numpy()
    ref = oracle.step(phi0, img.astype(np.float32), oracle_params(p))
    if mode == rsf_amd.MODE_LITERAL:
        assert np.array_equal(got, ref), float(np.abs(got - ref).max())
    else:
        err, _ = one_step_gate(got, ref)
        assert err <= 2e-5 * float(np.abs(ref).max()), err
    shapes = [shape, (P + 3, P + 8, P + 5)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    im = np.empty(total, np.uint8)
    ph = np.empty(total, np.float32)
    for k, ((off, w, h, l), shp) in enumerate(zip(blocks, shapes)):
        bi, bp, _ = scene(shp, tubes=3, seed=300 + P + k)
        im[off:off + bi.size] = bi.ravel()
        ph[off:off + bp.size] = bp.ravel()
    p2 = rsf_amd.make_params(sigma=sigma, epsilon=1.0, mode=mode, iterations=2)
    a, b = to_dev(ph), to_dev(ph)
    ctx.run_batched(blocks, to_dev(im), a, p2)
    ctx.run_sequential(blocks, to_dev(im), b, p2)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_config1_full_run_mask_gate(ctx, mode):
    """The CPU configuration at full size: 128^3 uint8, 40 tubes (seed 1), sigma 3,
    lambda 1, nu 65.025, mu 1, dt 0.1, eps 1, 100 iterations.  The sign masks of
    the device run and the oracle disagree on at most 1e-4 of the voxels."""
    n = 128
    img, _ = rsf_amd.synth_tubes((n, n, n), tubes=40, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=1)
    phi0 = rsf_amd.initial_phi(img, 1.0, 110.0)
    p = rsf_amd.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0, mode=mode,
                            iterations=100)
    print(f"config1 mode={mode}", flush=True)
    d_img, d_phi = to_dev(img), to_dev(phi0)
    ctx.run(d_img, d_phi, p)
    got = d_phi.cpu().numpy()
    ref = oracle.run(phi0, img.astype(np.float32), oracle_params(p), 100)
    dis = mask_disagreement(got, ref)
    if mode == rsf_amd.MODE_LITERAL:
        # bit-identical arithmetic: identical phi, not only the mask
        assert np.array_equal(got, ref), (dis, float(np.abs(got - ref).max()))
    else:
        # The explicit scheme at nu*dt = 6.5 amplifies rounding differences: the
        # oracle's own FP32 and FP64 runs of this case differ on 1.7e-4 of the
        # mask after 100 steps.  The fast path's regrouped sums stay within that
        # FP32-noise band.
        assert dis <= 4e-4, dis


@pytest.mark.parametrize("shape,phantom,iterations", [
    ((512, 512, 512), dict(tubes=300, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=2), 2),
    ((768, 1024, 1024), dict(tubes=2000, rmin=1.5, rmax=8.0, bias=20.0, noise=20.0, seed=5), 1),
], ids=["512cube", "1024x1024x768"])
def test_benchmark_volumes_literal_bit_identical(ctx, shape, phantom, iterations):
    """The volumes the benchmark times -- 512^3 (300 tubes, radius 2-6, +-20
    ramp, noise 15, seed 2) and 1024x1024x768 (2000 tubes, radius 1.5-8, noise
    20, seed 5), phi0 the signed distance of the thresholded sigma-1 blur --
    in LITERAL mode with the default kernels for their sizes (the larger one
    takes the long xy chunks): phi is identical to the oracle's bit for bit
    (the oracle runs multi-threaded)."""
    img, _ = rsf_amd.synth_tubes(shape, **phantom)
    phi0 = rsf_amd.initial_phi(img, 1.0, 110.0)
    p = rsf_amd.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0,
                            mode=rsf_amd.MODE_LITERAL, iterations=iterations)
    print(f"benchmark volume {shape} literal x{iterations}", flush=True)
    d_img, d_phi = to_dev(img), to_dev(phi0)
    ctx.run(d_img, d_phi, p)
    got = d_phi.cpu().numpy()
    del d_img, d_phi
    ref = oracle.run(phi0, img.astype(np.float32), oracle_params(p), iterations)
    assert np.array_equal(got, ref), float(np.abs(got - ref).max())


@pytest.mark.parametrize("mode", [rsf_amd.MODE_LITERAL, rsf_amd.MODE_FAST])
def test_volume_of_2_pow_30_voxels_matches_oracle_on_crops(ctx, mode):
    """1024^3 voxels (2^30: 4 GiB of phi) takes the kernels with 64-bit
    addressing -- the fused sweep 1 and the direct sweep 2 -- instead of the
    32-bit buffer-offset ones.  Two iterations are compared with the oracle on
    interior crops: a voxel's new phi depends only on phi and the image within
    12 voxels per iteration (10 for the blur, 2 for the curvature stencil), so at
    voxels 24 or more from the faces of a crop that itself lies 24 or more inside
    the volume (away from the pad band and the x-wrap), the oracle run on the
    crop alone computes exactly the whole-volume result.  LITERAL: bit for bit;
    FAST: the one-step-gate tolerance.  The crops include the far corner, beyond
    2^31 elements."""
    import torch

    n, T, c, mg = 1024, 2, 112, 24
    print(f"2^30 voxels mode={mode}", flush=True)
    ar = torch.arange(n, device="cuda", dtype=torch.float32)
    z, y, x = ar.view(n, 1, 1), ar.view(1, n, 1), ar.view(1, 1, n)
    f = torch.sin(x / 13.0) * torch.sin(y / 17.0 + 0.5) * torch.sin(z / 11.0 + 1.0)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(30)
    noise = torch.randn((n, n, n), device="cuda", generator=gen) * 15.0
    d_img = torch.clamp(torch.round(30.0 + 170.0 * (f > 0.3).float() + x * 0.02 + noise), 0, 255).to(torch.uint8)
    del noise
    d_phi = (8.0 * (0.3 - f)).contiguous()
    del f
    starts = [(mg, mg, mg), (n // 2 - c // 2, n // 2 - 40, n // 2 + 17), (n - mg - c, n - mg - c, n - mg - c)]
    crops = [(d_img[a:a + c, b:b + c, e:e + c].cpu().numpy(), d_phi[a:a + c, b:b + c, e:e + c].cpu().numpy())
             for a, b, e in starts]
    p = rsf_amd.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0, mode=mode,
                            iterations=T)
    ctx.run(d_img, d_phi, p)
    torch.cuda.synchronize()
    got = [d_phi[a + mg:a + c - mg, b + mg:b + c - mg, e + mg:e + c - mg].cpu().numpy() for a, b, e in starts]
    del d_img, d_phi
    torch.cuda.empty_cache()
    for (ci, cp), g in zip(crops, got):
        ref = oracle.run(cp, ci.astype(np.float32), oracle_params(p), T)[mg:c - mg, mg:c - mg, mg:c - mg]
        if mode == rsf_amd.MODE_LITERAL:
            assert np.array_equal(g, ref), float(np.abs(g - ref).max())
        else:
            err, tol = one_step_gate(g, ref)
            assert err <= tol, (err, tol)


def test_benchmark_volume_fast_one_step_gate(ctx):
    """FAST mode on the 512^3 benchmark volume: one step within the gate
    (max|dphi| <= 1e-4 max|phi|), and within 2e-5 max|phi| of the oracle."""
    n = 512
    img, _ = rsf_amd.synth_tubes((n, n, n), tubes=300, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=2)
    phi0 = rsf_amd.initial_phi(img, 1.0, 110.0)
    p = rsf_amd.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0,
                            mode=rsf_amd.MODE_FAST, iterations=1)
    print("benchmark volume 512^3 fast x1", flush=True)
    d_img, d_phi = to_dev(img), to_dev(phi0)
    ctx.run(d_img, d_phi, p)
    got = d_phi.cpu().numpy()
    del d_img, d_phi
    ref = oracle.run(phi0, img.astype(np.float32), oracle_params(p), 1)
    err, tol = one_step_gate(got, ref)
    assert err <= tol, (err, tol)
    assert err <= 2e-5 * float(np.abs(ref).max()), err


def test_nonfinite_phi_is_reported(ctx):
    """A NaN in phi0 spreads through the blurs; the run reports RSF_ERR_NONFINITE
    instead of returning garbage as a success (single volume and batched)."""
    import torch

    img, phi0, _ = scene((30, 32, 36), tubes=4, seed=81)
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=3)
    bad = phi0.copy()
    bad[15, 16, 18] = np.nan
    with pytest.raises(rsf_amd.RSFError) as e:
        ctx.run(to_dev(img), to_dev(bad), p)
    assert e.value.status == rsf_amd.ERR_NONFINITE and "NaN or infinite" in str(e.value)
    ctx.run(to_dev(img), to_dev(phi0), p)   # a clean run still succeeds afterwards
    shapes = [(30, 32, 36), (24, 26, 28)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    im = np.zeros(total, np.uint8)
    ph = np.zeros(total, np.float32)
    for (off, w, h, l), shp in zip(blocks, shapes):
        bi, bp, _ = scene(shp, tubes=3, seed=82)
        im[off:off + bi.size] = bi.ravel()
        ph[off:off + bp.size] = bp.ravel()
    ph[blocks[1][0] + 5] = np.inf
    with pytest.raises(rsf_amd.RSFError) as e:
        ctx.run_batched(blocks, to_dev(im), to_dev(ph), p)
    assert e.value.status == rsf_amd.ERR_NONFINITE


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_batched_at_other_tap_counts(ctx, mode):
    """sigma 2.5 (P = 17): the batched launch equals the blocks run one by one, and in
    literal mode each block equals the oracle's run on that block."""
    import torch

    shapes = [(30, 33, 36), (24, 40, 28), (36, 26, 32)]
    blocks, total = rsf_amd.pack_blocks(shapes)
    img = np.empty(total, np.uint8)
    phi = np.empty(total, np.float32)
    for k, ((off, w, h, l), shp) in enumerate(zip(blocks, shapes)):
        bi, bp, _ = scene(shp, tubes=3, seed=90 + k)
        img[off:off + bi.size] = bi.ravel()
        phi[off:off + bp.size] = bp.ravel()
    p = rsf_amd.make_params(sigma=2.5, epsilon=1.0, iterations=4, mode=mode)
    d_img = to_dev(img)
    a, b = to_dev(phi), to_dev(phi)
    ctx.run_batched(blocks, d_img, a, p)
    ctx.run_sequential(blocks, d_img, b, p)
    torch.cuda.synchronize()
    A = a.cpu().numpy()
    assert np.array_equal(A, b.cpu().numpy())
    if mode == rsf_amd.MODE_LITERAL:
        for (off, w, h, l), shp in zip(blocks, shapes):
            n = w * h * l
            ref = oracle.run(phi[off:off + n].reshape(shp), img[off:off + n].reshape(shp).astype(np.float32),
                             oracle_params(p), 4)
            assert np.array_equal(A[off:off + n].reshape(shp), ref)


@pytest.mark.parametrize("shape", [(36, 30, 52), (30, 33, 37)])
def test_literal_tiny_and_huge_gradients_take_the_exact_forms(ctx, shape):
    """Gradients below 2^-48 (|grad|^2 under the square root's scaling range)
    and above 2^64 (|grad|^2 overflows to infinity) are outside the fast
    forms' range: the literal interior flags those work items and the exact
    fix-up pass recomputes them.  One step equals the oracle bit for bit."""
    import torch

    img, phi0, _ = scene(shape, tubes=5, seed=91)
    l = shape[0]
    phi = phi0.astype(np.float64)
    phi[: l // 3] *= 1e-20            # tiny values: tiny differences
    phi[2 * l // 3:] *= 1e21          # huge values: |grad|^2 = inf in float
    phi = phi.astype(np.float32)
    assert np.isfinite(phi).all()
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, mode=rsf_amd.MODE_LITERAL, iterations=1)
    d_out = torch.empty(shape, dtype=torch.float32, device="cuda")
    ctx.step(to_dev(img), to_dev(phi), d_out, p)
    torch.cuda.synchronize()
    ref = oracle.step(phi, img.astype(np.float32), oracle_params(p))
    assert np.array_equal(d_out.cpu().numpy(), ref)


def test_literal_fixup_in_a_batched_launch(ctx):
    """The fix-up pass in a batched launch: of three uneven blocks, one has tiny
    phi values throughout, one huge ones in part, one ordinary ones; one step of
    each equals the oracle bit for bit."""
    shapes = [(30, 34, 44), (26, 40, 52), (33, 30, 37)]
    scales = [1e-20, None, 1.0]
    blocks, total = rsf_amd.pack_blocks(shapes)
    im = np.zeros(total, np.uint8)
    ph = np.zeros(total, np.float32)
    refs = []
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, mode=rsf_amd.MODE_LITERAL, iterations=1)
    for (off, w, h, l), shp, sc in zip(blocks, shapes, scales):
        bi, bp, _ = scene(shp, tubes=3, seed=93 + w)
        b = bp.astype(np.float64)
        if sc is None:
            b[: l // 2] *= 1e21
        else:
            b *= sc
        b = b.astype(np.float32)
        im[off:off + bi.size] = bi.ravel()
        ph[off:off + b.size] = b.ravel()
        refs.append(oracle.step(b, bi.astype(np.float32), oracle_params(p)))
    d_ph = to_dev(ph)
    ctx.run_batched(blocks, to_dev(im), d_ph, p)
    got = d_ph.cpu().numpy()
    for (off, w, h, l), ref in zip(blocks, refs):
        assert np.array_equal(got[off:off + w * h * l].reshape(l, h, w), ref)
