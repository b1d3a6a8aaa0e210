"""Streaming .npy I/O between files and device memory, and the file-to-file run
(CUDA_RSF/main.cpp:143-316): values equal numpy's, files larger than one
32 MiB staging chunk, and rsf_run_npy equal to the oracle."""
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


@pytest.mark.parametrize("dtype,shape", [
    (np.uint8, (80, 640, 700)),      # 35.8 MB: two chunks
    (np.float32, (40, 512, 520)),    # 42.6 MB: two chunks
    (np.float64, (9, 700, 1000)),    # converted per chunk
    (np.int16, (5, 6, 7)),
    (np.uint16, (5, 6, 7)),
    (np.bool_, (4, 5, 6)),
])
def test_npy_load_device_equals_numpy(ctx, tmp_path, dtype, shape):
    import torch

    rng = np.random.default_rng(3)
    if dtype == np.bool_:
        a = rng.random(shape) < 0.5
    elif np.issubdtype(dtype, np.integer):
        a = rng.integers(np.iinfo(dtype).min, np.iinfo(dtype).max, shape, endpoint=True).astype(dtype)
    else:
        a = (rng.normal(size=shape) * 100).astype(dtype)
    path = tmp_path / "a.npy"
    np.save(path, a)
    u8 = dtype in (np.uint8, np.bool_)
    d = torch.empty(shape, dtype=torch.uint8 if u8 else torch.float32, device="cuda")
    ctx.npy_load(str(path), d)
    torch.cuda.synchronize()
    ref = a.astype(np.uint8) if u8 else a.astype(np.float32)
    assert np.array_equal(d.cpu().numpy(), ref)


def test_npy_load_device_rejects_bad_destinations(ctx, tmp_path):
    import torch

    np.save(tmp_path / "f.npy", np.zeros((4, 4, 4), np.float32))
    with pytest.raises(rsf_amd.RSFError, match="not uint8"):
        ctx.npy_load(str(tmp_path / "f.npy"), torch.empty((4, 4, 4), dtype=torch.uint8, device="cuda"))
    with pytest.raises(rsf_amd.RSFError, match="too small"):
        ctx.npy_load(str(tmp_path / "f.npy"), torch.empty((4, 4, 3), device="cuda"))
    np.save(tmp_path / "flat.npy", np.zeros((4, 4), np.float32))
    with pytest.raises(rsf_amd.RSFError, match="3-D"):
        ctx.npy_load(str(tmp_path / "flat.npy"), torch.empty((4, 4, 4), device="cuda"))


def test_npy_save_device_round_trip(ctx, tmp_path):
    import torch

    shape = (33, 512, 520)   # 35 MB: two chunks
    src = torch.randn(shape, device="cuda")
    ctx.npy_save(str(tmp_path / "o.npy"), src)
    got = np.load(tmp_path / "o.npy")
    assert got.dtype == np.float32 and got.shape == shape
    assert np.array_equal(got, src.cpu().numpy())


@pytest.mark.parametrize("seed_is_phi0", [False, True])
def test_run_npy_equals_oracle(ctx, tmp_path, seed_is_phi0):
    shape = (28, 30, 34)
    img, phi_init, mask = scene(shape, tubes=5, seed=51)
    np.save(tmp_path / "img.npy", img)
    if seed_is_phi0:
        np.save(tmp_path / "seed.npy", phi_init)
        phi0 = phi_init
    else:
        seed = (mask * 255).astype(np.uint8)
        np.save(tmp_path / "seed.npy", seed)
        phi0 = rsf_amd.sdf_from_seed(seed.astype(np.float32), 13.0)
    p = rsf_amd.make_params(sigma=1.0, epsilon=1.0, mode=rsf_amd.MODE_LITERAL, iterations=3)
    ms = ctx.run_npy(str(tmp_path / "img.npy"), str(tmp_path / "seed.npy"), str(tmp_path / "phi.npy"), p,
                     seed_is_phi0=seed_is_phi0, per_iteration=True)
    assert ms.shape == (3,) and np.all(ms > 0)
    ref = oracle.run(phi0, img.astype(np.float32), oracle_params(p), 3)
    assert np.array_equal(np.load(tmp_path / "phi.npy"), ref)
