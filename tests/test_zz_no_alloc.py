"""The allocation-free step (rsf_step_device_noalloc) against rsf_step_device:
same bits, called through the C ABI twice in a row (the second volume smaller,
with rows that are not 16-byte aligned), and its limits reported as errors."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rsf_amd
from gpu_util import oracle_params, scene, to_dev

pytestmark = pytest.mark.gpu


def _noalloc(ctx, img, phi_in, phi_out, p):
    l, h, w = phi_in.shape
    return rsf_amd.lib().rsf_step_device_noalloc(ctx.handle, img.data_ptr(), rsf_amd._dtype_code(img),
                                                 phi_in.data_ptr(), phi_out.data_ptr(), w, h, l, C.byref(p))


def test_noalloc_step_equals_step_device_twice_in_a_row():
    import torch

    ctx = rsf_amd.Context(0)
    try:
        p = rsf_amd.make_params(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0,
                                mode=rsf_amd.MODE_LITERAL, iterations=1)
        for shape, dtype in (((100, 110, 124), np.uint8), ((40, 44, 53), np.float32)):
            print(f"noalloc step {shape} {np.dtype(dtype).name}", flush=True)
            img, phi0, _ = scene(shape, tubes=8, seed=sum(shape))
            d_img, d_phi = to_dev(img.astype(dtype)), to_dev(phi0)
            a, b = torch.full_like(d_phi, np.nan), torch.full_like(d_phi, np.nan)
            torch.cuda.synchronize()
            rsf_amd.check(_noalloc(ctx, d_img, d_phi, a, p))
            ctx.step(d_img, d_phi, b, p)
            torch.cuda.synchronize()
            got = a.cpu().numpy()
            assert np.array_equal(got, b.cpu().numpy())
            assert np.array_equal(got, oracle.step(phi0, img.astype(np.float32), oracle_params(p)))
    finally:
        ctx.close()


def test_noalloc_step_rejects_what_it_cannot_do():
    import torch

    ctx = rsf_amd.Context(0)
    try:
        small = torch.zeros((32, 32, 32), device="cuda")
        img = torch.zeros((32, 32, 32), dtype=torch.uint8, device="cuda")
        L = rsf_amd.lib()
        p = rsf_amd.make_params(sigma=3.0, mode=rsf_amd.MODE_LITERAL)
        # a volume above the scratch's 2^23 voxels: refused before any device work
        st = L.rsf_step_device_noalloc(ctx.handle, img.data_ptr(), rsf_amd.DTYPE_U8, small.data_ptr(),
                                       small.data_ptr() + (64 << 20), 256, 256, 129, C.byref(p))
        assert st == rsf_amd.ERR_INVALID_ARGUMENT and b"scratch" in L.rsf_last_error()
        pf = rsf_amd.make_params(sigma=3.0, mode=rsf_amd.MODE_FAST)
        out = torch.empty_like(small)
        assert _noalloc(ctx, img, small, out, pf) == rsf_amd.ERR_UNSUPPORTED
        assert _noalloc(ctx, img, small, small, p) == rsf_amd.ERR_INVALID_ARGUMENT
    finally:
        ctx.close()
