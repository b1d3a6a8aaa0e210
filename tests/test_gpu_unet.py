"""The U-Net seed producer on the GPU: the device vote bit-identical to the
oracle's, the whole segment() on the GPU against the model on the CPU voted by
the oracle, the seed hand-off into the device SDF and the RSF run, and
gradients from an RSF loss back into the U-Net."""
import numpy as np
import pytest

import rsf_amd
from gpu_util import scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch  # noqa: F401

    c = rsf_amd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("shape", [(40, 100, 130), (17, 65, 97), (16, 64, 64)])
def test_vote_on_device_equals_oracle(shape):
    import torch

    from oracle import unet_vote
    from rsf_amd import unet

    sl = unet.tile_slices(shape)
    pred = torch.rand((len(sl), unet.NUM_PICS, unet.HEIGHT, unet.WIDTH), generator=torch.Generator().manual_seed(1))
    pc, mc = unet_vote.tile_vote(pred, shape, sl)
    pg, mg = unet.vote(pred.cuda(), shape, sl)
    assert torch.equal(pg.cpu(), pc) and torch.equal(mg.cpu(), mc)


def test_segment_on_device_matches_cpu_model_and_oracle_vote():
    import torch

    from oracle import unet_vote
    from rsf_amd import unet

    torch.manual_seed(0)
    model = unet.UNet3D().eval()
    img, _, _ = scene((32, 96, 96), tubes=5, seed=61)
    sl = unet.tile_slices(img.shape)
    pred = unet.predict_tiles(model, torch.from_numpy(unet.normalize(img)), sl, batch=8)
    pc, _ = unet_vote.tile_vote(pred, img.shape, sl)
    pg, _ = unet.segment(model.cuda(), img, device="cuda", batch=32)
    # the model runs in FP32 on both sides (different convolution kernels)
    assert float((pg.cpu() - pc).abs().max()) < 1e-4


def test_seed_to_phi0_to_rsf(ctx):
    import torch

    from rsf_amd import unet

    torch.manual_seed(0)
    model = unet.UNet3D().cuda().eval()
    img, _, _ = scene((32, 64, 64), tubes=4, seed=62)
    prob, mask = unet.segment(model, img, device="cuda")
    seed = unet.seed_volume(mask).contiguous()
    phi = torch.empty(img.shape, device="cuda")
    ctx.sdf_from_seed(seed, phi, 13.0)
    ref = rsf_amd.sdf_from_seed(seed.cpu().numpy().astype(np.float32), 13.0)
    assert np.array_equal(phi.cpu().numpy(), ref)
    p = rsf_amd.make_params(sigma=1.0, epsilon=1.0, iterations=3)
    ctx.run(torch.from_numpy(img).cuda(), phi, p)
    assert bool(torch.isfinite(phi).all())


def test_rsf_loss_backpropagates_into_the_unet():
    import torch

    from rsf_amd import unet
    from rsf_amd.autograd import bce_inside_loss, rsf_unrolled

    torch.manual_seed(0)
    model = unet.UNet3D().cuda()
    img, _, mask = scene((16, 64, 64), tubes=4, seed=63)
    x = torch.from_numpy(unet.normalize(img)).float().cuda()[None, None]
    prob = model(x)[0, 0]
    phi0 = unet.phi0_from_probability(prob, scale=2.0).float().contiguous()
    l, h, w = img.shape
    out = rsf_unrolled(phi0.reshape(-1), torch.tensor([1.0, 1.0, 65.025, 1.0], device="cuda"),
                       torch.from_numpy(img).cuda().reshape(-1), [(0, w, h, l)], K=4, checkpoint_every=2,
                       sigma=1.0)
    loss = bce_inside_loss(out, torch.from_numpy(mask).cuda().reshape(-1).float())
    loss.backward()
    grads = [p.grad for p in model.parameters()]
    assert all(g is not None and bool(torch.isfinite(g).all()) for g in grads)
    assert sum(float(g.abs().sum()) for g in grads) > 0


@pytest.mark.parametrize("dtype", [np.uint8, np.float32, np.uint16, np.int16])
def test_normalize_on_device_equals_host(dtype):
    import torch

    from rsf_amd import unet

    rng = np.random.default_rng(3)
    vol = (rng.random((20, 70, 90)) * 250 + 3).astype(dtype)
    got = unet.normalize_device(vol)
    ref = unet.normalize(vol)
    assert got.dtype == torch.from_numpy(ref).dtype
    assert torch.equal(got.cpu(), torch.from_numpy(ref))
    assert torch.equal(unet.normalize_device(vol, "blur").cpu(), torch.from_numpy(unet.normalize(vol, "blur")))
