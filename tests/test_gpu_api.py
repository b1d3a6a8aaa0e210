"""GPU checks of the C ABI's caching and argument rules: the held-image
preparation is never served for another image, captured graphs are keyed by
the parameters' exact bits, overlapping phi ranges are refused, the device
wrappers refuse host tensors, and rsf_workspace_size matches what a run
allocates."""
import ctypes as C

import numpy as np
import pytest

import rsf_amd
from gpu_util import scene, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch  # noqa: F401

    c = rsf_amd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("kw", [dict(mode=rsf_amd.MODE_FAST), dict(mode=rsf_amd.MODE_LITERAL, sigma2=1.0)],
                         ids=["fast", "literal_sigma2"])
def test_held_image_preparation_is_not_served_for_another_image(ctx, kw):
    """hold(X); step(X); step(Y); step(X): the last step must use X's float copy
    and blur, not Y's (which the step on Y wrote into the shared workspace)."""
    import torch

    shape = (36, 38, 40)
    X, phi0, _ = scene(shape, tubes=5, seed=11)
    Y, _, _ = scene(shape, tubes=9, seed=12)
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=1, **kw)
    dX, dY, dphi = to_dev(X), to_dev(Y), to_dev(phi0)
    out = torch.empty_like(dphi)
    ctx.hold_image(dX)
    try:
        ctx.step(dX, dphi, out, p)
        ctx.step(dY, dphi, out, p)
        ctx.step(dX, dphi, out, p)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        ctx.run(dX, dphi.clone(), rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=2, **kw))
        ctx.step(dX, dphi, out, p)
        got_after_run = out.cpu().numpy()
    finally:
        ctx.release_image()
    fresh = rsf_amd.Context(0)
    ref = torch.empty_like(dphi)
    fresh.step(dX, dphi, ref, p)
    torch.cuda.synchronize()
    fresh.close()
    assert np.array_equal(got, ref.cpu().numpy())
    assert np.array_equal(got_after_run, ref.cpu().numpy())


def test_graph_cache_keys_parameters_by_their_bits(ctx):
    """Two runs whose nu differs by one ulp (both print as 65.025 at six digits)
    must not share a captured graph: the second equals a fresh context's run."""
    import torch

    img, phi0, _ = scene((40, 42, 44), tubes=6, seed=13)
    nu0 = np.float32(65.025)
    nu1 = np.nextafter(nu0, np.float32(100))
    assert f"{float(nu0):g}" == f"{float(nu1):g}"
    d_img = to_dev(img)
    outs = []
    for nu in (nu0, nu1):
        d = to_dev(phi0)
        ctx.run(d_img, d, rsf_amd.make_params(sigma=3.0, epsilon=1.0, nu=float(nu), iterations=6))
        outs.append(d.cpu().numpy())
    fresh = rsf_amd.Context(0)
    d = to_dev(phi0)
    fresh.run(d_img, d, rsf_amd.make_params(sigma=3.0, epsilon=1.0, nu=float(nu1), iterations=6))
    torch.cuda.synchronize()
    fresh.close()
    assert np.array_equal(outs[1], d.cpu().numpy())
    assert not np.array_equal(outs[0], outs[1])


def test_overlapping_phi_ranges_are_refused(ctx):
    """rsf_step_blocks_device and rsf_step_device refuse phi_out ranges that
    overlap phi_in (sweep 2 would read neighbours other workgroups overwrite);
    the adjoint refuses a grad_in overlapping grad_out."""
    import torch

    shape = (30, 32, 36)
    img, phi0, _ = scene(shape, tubes=3, seed=14)
    n = phi0.size
    base = torch.zeros(n + 8, dtype=torch.float32, device="cuda")
    base[:n] = to_dev(phi0).view(-1)
    d_img = to_dev(img)
    L = rsf_amd.lib()
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=1)
    blk = rsf_amd.blocks_array([(0, 36, 32, 30)])
    st = L.rsf_step_blocks_device(ctx.handle, blk, 1, d_img.data_ptr(), rsf_amd.DTYPE_U8, base.data_ptr(),
                                  base.data_ptr() + 16, C.byref(p))
    assert st == rsf_amd.ERR_INVALID_ARGUMENT and b"overlap" in L.rsf_last_error()
    st = L.rsf_step_device(ctx.handle, d_img.data_ptr(), rsf_amd.DTYPE_U8, base.data_ptr(), base.data_ptr() + 16,
                           36, 32, 30, C.byref(p))
    assert st == rsf_amd.ERR_INVALID_ARGUMENT
    g = torch.zeros(n + 8, device="cuda")
    gp = torch.zeros(4, device="cuda")
    st = L.rsf_step_adjoint_device(ctx.handle, blk, 1, d_img.data_ptr(), rsf_amd.DTYPE_U8, base.data_ptr(),
                                   g.data_ptr(), g.data_ptr() + 32, gp.data_ptr(),
                                   C.byref(rsf_amd.make_params(sigma=3.0, epsilon=1.0, mode=rsf_amd.MODE_FAST)))
    assert st == rsf_amd.ERR_INVALID_ARGUMENT
    # disjoint ranges next to each other are fine
    out = torch.empty(n, device="cuda")
    ctx.step(d_img, base[:n].view(shape), out.view(shape), p)


def test_device_wrappers_refuse_host_tensors(ctx):
    import torch

    img, phi0, _ = scene((24, 24, 24), tubes=2, seed=15)
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=1)
    d_img, d_phi = to_dev(img), to_dev(phi0)
    with pytest.raises(ValueError):
        ctx.step(torch.from_numpy(img), d_phi, torch.empty_like(d_phi), p)
    with pytest.raises(ValueError):
        ctx.run(d_img, torch.from_numpy(phi0), p)
    with pytest.raises(TypeError):
        ctx.run(d_img, d_phi.double(), p)
    with pytest.raises(ValueError):
        ctx.run_batched([(0, 24, 24, 24), (24 ** 3, 24, 24, 24)], d_img.view(-1), d_phi.view(-1), p)
    with pytest.raises(ValueError):
        ctx.step_blocks([(0, 24, 24, 24)], d_img.view(-1), d_phi.view(-1), torch.empty(100, device="cuda"), p)


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
def test_workspace_size_matches_the_allocation(mode):
    """A fresh context's first plain run allocates what rsf_workspace_size
    reports (plus tile lists, flags and graph storage: under 64 MiB)."""
    import torch

    shape = (180, 200, 256)
    l, h, w = shape
    img = torch.full(shape, 40, dtype=torch.uint8, device="cuda")
    phi = torch.linspace(-5, 5, w, device="cuda").expand(shape).contiguous()
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=2, mode=mode)
    b = C.c_int64()
    rsf_amd.check(rsf_amd.lib().rsf_workspace_size(C.byref(p), w, h, l, C.byref(b)))
    torch.cuda.synchronize()
    c = rsf_amd.Context(0)
    free0 = torch.cuda.mem_get_info()[0]
    c.run(img, phi, p)
    torch.cuda.synchronize()
    used = free0 - torch.cuda.mem_get_info()[0]
    c.close()
    assert b.value <= used <= b.value + (64 << 20), (b.value, used)
