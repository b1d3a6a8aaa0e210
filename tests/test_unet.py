"""U-Net seed producer on the CPU: the tile list and the oracle's vote against a
literal restatement of Unet/run_model.py's loop, the model's shape contract,
the identity-model check of tiling + prediction + vote, and the product vote's
refusal of host tensors (it runs on the device only)."""
import math

import numpy as np
import pytest
import torch

from oracle import unet_vote
from rsf_amd import unet


def reference_tiles(shape):
    """run_model.py:39-46,72-85 as written: the slices it visits, in order."""
    l, h, w = shape
    step = int(unet.HEIGHT / unet.STEP_RATIO)
    out = []
    for i in range(int(np.ceil(l / unet.NUM_PICS))):
        for j in range(int(np.ceil(h / step))):
            for k in range(int(np.ceil(w / step))):
                out.append((min(i * unet.NUM_PICS, l - unet.NUM_PICS), min((i + 1) * unet.NUM_PICS, l),
                            min(j * step, h - unet.HEIGHT), min(j * step + unet.HEIGHT, h),
                            min(k * step, w - unet.WIDTH), min(k * step + unet.WIDTH, w)))
    return out


def reference_vote(pred, shape, slices):
    """run_model.py:29-35,86-91: float64 count and total, tile by tile."""
    count = np.zeros(shape, np.float64)
    total = np.zeros(shape, np.float64)
    for t, (z0, z1, y0, y1, x0, x1) in enumerate(slices):
        count[z0:z1, y0:y1, x0:x1] = count[z0:z1, y0:y1, x0:x1] + pred[t]
        total[z0:z1, y0:y1, x0:x1] += 1
    prob = np.zeros(shape, np.float64)
    m = total != 0
    prob[m] = count[m] / total[m]
    return prob, prob >= 0.5


class _Identity(torch.nn.Module):
    def forward(self, x):
        return x


SHAPES = [(16, 64, 64), (32, 128, 128), (40, 100, 130), (17, 65, 97)]


@pytest.mark.parametrize("shape", SHAPES)
def test_tile_slices_follow_the_reference_loop(shape):
    got = unet.tile_slices(shape)
    assert got == reference_tiles(shape)
    for (z0, z1, y0, y1, x0, x1) in got:
        assert (z1 - z0, y1 - y0, x1 - x0) == (unet.NUM_PICS, unet.HEIGHT, unet.WIDTH)


def test_tile_slices_reject_volumes_smaller_than_a_tile():
    with pytest.raises(ValueError):
        unet.tile_slices((15, 64, 64))


@pytest.mark.parametrize("shape", SHAPES)
def test_oracle_vote_equals_reference_bit_for_bit(shape):
    sl = unet.tile_slices(shape)
    rng = np.random.default_rng(len(sl))
    pred = rng.random((len(sl), unet.NUM_PICS, unet.HEIGHT, unet.WIDTH), dtype=np.float32)
    pred[0, 0, 0, :4] = [1e-30, 3e-12, 0.5, 1.0]     # values whose float64 sums round
    prob, mask = unet_vote.tile_vote(torch.from_numpy(pred), shape, sl)
    rp, rm = reference_vote(pred, shape, sl)
    assert np.array_equal(prob.numpy(), rp)
    assert np.array_equal(mask.numpy(), rm)


def test_product_vote_and_segment_refuse_the_host():
    shape = (16, 64, 64)
    sl = unet.tile_slices(shape)
    with pytest.raises(ValueError, match="device"):
        unet.vote(torch.zeros((len(sl), unet.NUM_PICS, unet.HEIGHT, unet.WIDTH)), shape, sl)
    with pytest.raises(ValueError, match="HIP device"):
        unet.segment(_Identity(), np.zeros(shape, np.uint8), device="cpu")


def test_model_shape_and_range():
    m = unet.UNet3D().eval()
    with torch.no_grad():
        y = m(torch.rand(2, 1, unet.NUM_PICS, unet.HEIGHT, unet.WIDTH))
    assert y.shape == (2, 1, unet.NUM_PICS, unet.HEIGHT, unet.WIDTH)
    assert float(y.min()) > 0 and float(y.max()) < 1


def test_load_model_reads_a_tensor_state_dict(tmp_path):
    m = unet.UNet3D()
    torch.save(m.state_dict(), tmp_path / "w.pth")
    m2 = unet.load_model(str(tmp_path / "w.pth"))
    for a, b in zip(m.state_dict().values(), m2.state_dict().values()):
        assert torch.equal(a, b)


def test_tiles_of_an_identity_model_vote_back_to_the_normalized_volume():
    rng = np.random.default_rng(4)
    vol = rng.integers(0, 256, (20, 70, 90), dtype=np.uint8)
    x = unet.normalize(vol)
    sl = unet.tile_slices(vol.shape)
    pred = unet.predict_tiles(_Identity(), torch.from_numpy(x), sl, batch=7)
    prob, mask = unet_vote.tile_vote(pred, vol.shape, sl)
    # every voxel averages n copies of float32(x): the float64 mean is float32(x) exactly
    assert np.array_equal(prob.numpy(), x.astype(np.float32).astype(np.float64))
    assert np.array_equal(mask.numpy(), x.astype(np.float32) >= 0.5)


@pytest.mark.parametrize("mode", [None, "blur", "pff"])
def test_normalize_modes(mode):
    rng = np.random.default_rng(5)
    vol = rng.integers(1, 256, (4, 30, 30), dtype=np.uint8)
    x = unet.normalize(vol, mode)
    assert x.dtype == np.float64 and x.shape == vol.shape
    assert x.min() == 0.0 and x.max() == 1.0
    if mode is None:
        assert np.array_equal(np.argsort(x.ravel(), kind="stable"), np.argsort(vol.ravel(), kind="stable"))


def test_phi0_from_probability_sign():
    p = torch.tensor([0.1, 0.5, 0.9], dtype=torch.float64)
    phi = unet.phi0_from_probability(p, scale=2.0)
    assert phi[0] > 0 and phi[1] == 0 and phi[2] < 0
    assert math.isclose(float(phi[2]), -2.0 * math.log(9.0), rel_tol=1e-12)
