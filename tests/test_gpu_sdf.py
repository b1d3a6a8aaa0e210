"""phi0 = sdf(seed - 13) (CUDA_RSF/main.cpp:174-183) on the device: bit-identical
to the host form (which the CPU suite checks against scipy's exact EDT) below a
4096-voxel diagonal, and equal to scipy's EDT rounded once beyond it."""
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
    # an unknown dtype code, an empty dimension and NULL buffers: rejected before any access
    assert lib.rsf_sdf_from_seed_device(ctx.handle, seed.data_ptr(), 7, d.data_ptr(), 4, 4, 4, 13.0) != 0
    assert lib.rsf_sdf_from_seed_device(ctx.handle, seed.data_ptr(), rsf_amd.DTYPE_F32, d.data_ptr(),
                                        0, 4, 4, 13.0) != 0
    assert lib.rsf_sdf_from_seed_device(ctx.handle, None, rsf_amd.DTYPE_F32, d.data_ptr(), 4, 4, 4, 13.0) != 0
    assert lib.rsf_sdf_from_seed_host(ctx.handle, None, None, 4, 4, 4, 13.0) != 0


def _edt_phi(seed, shift=13.0):
    """The exact field from scipy's float64 EDT: positive inside {seed - shift > 0},
    distances capped at the volume diagonal, minus 1/2, rounded once to float."""
    from scipy import ndimage

    inside = (seed.astype(np.float32) - np.float32(shift)) > 0
    l, h, w = seed.shape
    cap = np.sqrt(float(w) * w + float(h) * h + float(l) * l)
    d_in = ndimage.distance_transform_edt(inside)
    d_out = ndimage.distance_transform_edt(~inside)
    return np.where(inside, np.minimum(d_in, cap) - 0.5, -(np.minimum(d_out, cap) - 0.5)).astype(np.float32)


@pytest.mark.parametrize("shape", [(4, 40, 4100), (4100, 6, 5), (3, 2900, 2900)])
def test_device_sdf_beyond_4096_diagonal_is_exact(ctx, shape):
    """Volumes whose diagonal reaches 4096 voxels carry the squared distances in
    double: the field equals scipy's exact EDT rounded once to float."""
    import torch

    rng = np.random.default_rng(sum(shape))
    seed = np.zeros(shape, np.float32)
    for _ in range(5):
        c = [int(rng.integers(0, n)) for n in shape]
        seed[tuple(slice(max(0, ci - 2), ci + 3) for ci in c)] = 255.0
    l, h, w = shape
    assert w * w + h * h + l * l >= 4096 ** 2
    d = torch.empty(shape, dtype=torch.float32, device="cuda")
    ctx.sdf_from_seed(to_dev(seed), d, 13.0)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), _edt_phi(seed))
