"""phi0 = sdf(seed - 13) (CUDA_RSF/main.cpp:174-183) on the device: bit-identical
to the host form, which the CPU suite checks against scipy's exact EDT."""
import numpy as np
import pytest

import rsf_amd
from gpu_util import to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch  # noqa: F401

    c = rsf_amd.Context(0)
    yield c
    c.close()


def _blobs(shape, seed, count=6, rmax=6.0):
    rng = np.random.default_rng(seed)
    l, h, w = shape
    z, y, x = np.meshgrid(np.arange(l), np.arange(h), np.arange(w), indexing="ij")
    s = np.zeros(shape, np.float32)
    for _ in range(count):
        c = rng.uniform([0, 0, 0], [l, h, w])
        r = rng.uniform(1.0, rmax)
        s[(z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2 <= r * r] = 255.0
    return s


def _cases():
    out = []
    for k, shape in enumerate([(20, 22, 24), (33, 17, 70), (64, 64, 64), (9, 70, 5), (1, 1, 7), (40, 1, 33)]):
        out.append((f"blobs{k}", _blobs(shape, k)))
    s = np.zeros((30, 31, 65), np.float32)
    s[17, 4, 60] = 200.0
    out.append(("one_voxel", s))
    out.append(("empty", np.zeros((12, 13, 14), np.float32)))
    out.append(("full", np.full((12, 13, 14), 255.0, np.float32)))
    s = np.zeros((40, 24, 28), np.float32)
    s[3:6, :, :] = 255.0           # a slab: most planes hold one kind only
    s[30, 10, :] = 14.0            # just above the shift: inside
    s[20, :, 5] = 13.0             # exactly the shift: outside
    out.append(("slab_line", s))
    rng = np.random.default_rng(9)
    out.append(("noise", (rng.random((25, 26, 27)) < 0.02).astype(np.float32) * 255))
    return out


@pytest.mark.parametrize("name,seed", _cases(), ids=[c[0] for c in _cases()])
@pytest.mark.parametrize("dtype", [np.float32, np.uint8])
def test_device_sdf_equals_host(ctx, name, seed, dtype):
    import torch

    ref = rsf_amd.sdf_from_seed(seed, 13.0)
    d = torch.empty(seed.shape, dtype=torch.float32, device="cuda")
    ctx.sdf_from_seed(to_dev(seed.astype(dtype)), d, 13.0)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    assert np.array_equal(got, ref), (name, float(np.abs(got - ref).max()))


def test_device_sdf_rejects_bad_input(ctx):
    import torch

    d = torch.empty((4, 4, 4), device="cuda")
    with pytest.raises(TypeError):
        ctx.sdf_from_seed(torch.zeros((4, 4, 4), dtype=torch.int32, device="cuda"), d)
    seed = torch.zeros((4, 4, 4), device="cuda")
    lib = rsf_amd.lib()
    # an unknown dtype code, and a diagonal of 4096 voxels or more: rejected before any access
    assert lib.rsf_sdf_from_seed_device(ctx.handle, seed.data_ptr(), 7, d.data_ptr(), 4, 4, 4, 13.0) != 0
    assert lib.rsf_sdf_from_seed_device(ctx.handle, seed.data_ptr(), rsf_amd.DTYPE_F32, d.data_ptr(),
                                        4096, 1, 1, 13.0) != 0
