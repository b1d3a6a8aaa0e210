"""U-Net seed producer on PyTorch-ROCm: tiled inference with averaged votes.

The reference segments an initial contour with a 3-D U-Net applied to
16 x 64 x 64 tiles (Unet/config.py:3-7) at a stride of 16 along z and 32 along
y and x, and averages the tiles' outputs voxel by voxel (Unet/run_model.py:
21-91).  Its model definition (model_structures/UNET_3d_2l.py, `unetseg.py` in
README.md) is not in the repository, so `UNet3D` here is a two-level 3-D U-Net
of this package's own design with the same interface: one input channel, one
sigmoid output channel, shape-preserving on 16 x 64 x 64 tiles.  State dicts
are loaded with `torch.load(..., weights_only=True)`.

What follows the reference exactly:
* `normalize`: the three input preparations of run_model.py:53-66;
* `tile_slices`: the tile list of run_model.py:39-46,72-84 in loop order,
  including the repeated edge tiles that its clamping produces;
* `vote`: count / total in float64 with every voxel's tile outputs summed in
  the loop's order (run_model.py:86-91), then probability >= 0.5
  (run_model.py:29-35), on the device.  The reference unpacks that function's results into
  swapped names (run_model.py:91): the array it saves as "processed" is the
  probability and the one it saves as "probability" is the 0/1 map.

The hand-off to the RSF engine: `seed_volume` is the 0/1 map scaled to
0/255 uint8, so that phi0 = sdf(seed - 13) (CUDA_RSF/main.cpp:174-183) is
the signed distance of the >= 0.5 region; `phi0_from_probability` is a
differentiable alternative (negative inside, the convention of
`autograd.bce_inside_loss`) for training the U-Net through the unrolled RSF
steps (`autograd.rsf_unrolled`).
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn as nn

NUM_PICS, HEIGHT, WIDTH = 16, 64, 64   # Unet/config.py:3-7
STEP_RATIO = 2                          # run_model.py:72


# ------------------------------------------------------------------- model
def _block(cin, cout):
    return nn.Sequential(
        nn.Conv3d(cin, cout, 3, padding=1), nn.BatchNorm3d(cout), nn.ReLU(inplace=True),
        nn.Conv3d(cout, cout, 3, padding=1), nn.BatchNorm3d(cout), nn.ReLU(inplace=True))


class UNet3D(nn.Module):
    """Two-level 3-D U-Net (encoder 16 -> 32 channels, bottleneck 64, decoder with
    skip connections), sigmoid output.  Input (N, 1, D, H, W) with D, H, W
    divisible by 4."""

    def __init__(self, base: int = 16):
        super().__init__()
        self.enc1 = _block(1, base)
        self.enc2 = _block(base, 2 * base)
        self.mid = _block(2 * base, 4 * base)
        self.pool = nn.MaxPool3d(2)
        self.up2 = nn.ConvTranspose3d(4 * base, 2 * base, 2, stride=2)
        self.dec2 = _block(4 * base, 2 * base)
        self.up1 = nn.ConvTranspose3d(2 * base, base, 2, stride=2)
        self.dec1 = _block(2 * base, base)
        self.head = nn.Conv3d(base, 1, 1)

    def forward(self, x):
        e1 = self.enc1(x)
        e2 = self.enc2(self.pool(e1))
        m = self.mid(self.pool(e2))
        d2 = self.dec2(torch.cat([self.up2(m), e2], 1))
        d1 = self.dec1(torch.cat([self.up1(d2), e1], 1))
        return torch.sigmoid(self.head(d1))


def load_model(path: str | None = None, device="cpu") -> UNet3D:
    """UNet3D in eval mode, with a state dict from `path` (tensors only) if given."""
    m = UNet3D()
    if path is not None:
        m.load_state_dict(torch.load(path, map_location="cpu", weights_only=True))
    return m.to(device).eval()


# ------------------------------------------------------------ preparation
def normalize(volume: np.ndarray, mode: str | None = None) -> np.ndarray:
    """The model input of run_model.py:53-66, float64 (l, h, w) in [0, 1].

    mode None: min-max scaling; "blur": each z-slice Gaussian-blurred (sigma 1,
    scipy, the slice's dtype kept) first; "pff": floor(slice / blur * 255),
    min-max scaled, then scaled again."""
    from scipy.ndimage import gaussian_filter

    if volume.ndim != 3:
        raise ValueError("volume must be (Z, Y, X)")
    if mode is None:
        x = volume
    elif mode == "blur":
        x = np.asarray([gaussian_filter(volume[i], sigma=1) for i in range(volume.shape[0])])
    elif mode == "pff":
        blur = [gaussian_filter(volume[i], sigma=1) for i in range(volume.shape[0])]
        with np.errstate(divide="ignore", invalid="ignore"):
            pff = [np.floor(volume[i] / blur[i] * 255) for i in range(volume.shape[0])]
        x = (pff - np.min(pff)) / (np.max(pff) - np.min(pff))
    else:
        raise ValueError(f"unknown mode {mode!r}")
    x = np.asarray(x)
    return (x - np.min(x)) / (np.max(x) - np.min(x))


def normalize_device(volume: np.ndarray, mode: str | None = None, device="cuda") -> torch.Tensor:
    """`normalize` with the result on `device`.  Mode None (min-max scaling, the
    default) runs there with numpy's arithmetic -- integer volumes subtract the
    minimum in their own type and divide in float64, float volumes stay in their
    type -- so the values are normalize()'s bit for bit; the scipy modes run on
    the host as in the reference and are uploaded."""
    if mode is not None or volume.dtype not in (np.uint8, np.uint16, np.float32, np.float64):
        # (signed integers: numpy's x - min may wrap; kept on the host to match it)
        return torch.from_numpy(normalize(volume, mode)).to(device)
    if volume.ndim != 3:
        raise ValueError("volume must be (Z, Y, X)")
    t = torch.from_numpy(np.ascontiguousarray(volume)).to(device)
    if not t.dtype.is_floating_point:
        t = t.long()   # unsigned: x - min is exact in the source type and in int64 alike
    mn, mx = t.min(), t.max()
    if t.dtype.is_floating_point:
        return (t - mn) / (mx - mn)
    return (t - mn).double() / (mx - mn).double()


def tile_slices(shape):
    """(z0, z1, y0, y1, x0, x1) of every tile of a (l, h, w) volume in the order
    run_model.py:82-85 visits them.  Each axis must hold at least one tile."""
    l, h, w = shape
    if l < NUM_PICS or h < HEIGHT or w < WIDTH:
        raise ValueError(f"volume {shape} is smaller than one {NUM_PICS}x{HEIGHT}x{WIDTH} tile")
    step = HEIGHT // STEP_RATIO
    zs = np.arange(math.ceil(l / NUM_PICS))
    ys = np.arange(math.ceil(h / step))
    xs = np.arange(math.ceil(w / step))
    z0 = np.minimum(zs * NUM_PICS, l - NUM_PICS)
    z1 = np.minimum((zs + 1) * NUM_PICS, l)
    y0 = np.minimum(ys * step, h - HEIGHT)
    y1 = np.minimum(ys * step + HEIGHT, h)
    x0 = np.minimum(xs * step, w - WIDTH)
    x1 = np.minimum(xs * step + WIDTH, w)
    return [(int(z0[i]), int(z1[i]), int(y0[j]), int(y1[j]), int(x0[k]), int(x1[k]))
            for i in range(len(zs)) for j in range(len(ys)) for k in range(len(xs))]


# ---------------------------------------------------------------- inference
def predict_tiles(model, x: torch.Tensor, slices, batch: int = 64) -> torch.Tensor:
    """Model outputs of every tile, (T, 16, 64, 64) float32, in tile order.
    x: normalized (l, h, w) volume on the model's device."""
    out = torch.empty((len(slices), NUM_PICS, HEIGHT, WIDTH), dtype=torch.float32, device=x.device)
    xf = x.float()
    with torch.no_grad():
        for b0 in range(0, len(slices), batch):
            sl = slices[b0:b0 + batch]
            tiles = torch.stack([xf[z0:z1, y0:y1, x0:x1] for (z0, z1, y0, y1, x0, x1) in sl]).unsqueeze(1)
            out[b0:b0 + len(sl)] = model(tiles)[:, 0]
    return out


def vote(pred: torch.Tensor, shape, slices):
    """count / total of run_model.py:86-91 from the tile outputs `pred` (T, 16, 64, 64)
    on a HIP device (rsf_tile_vote_device).

    Every voxel sums its covering tiles' outputs in float64 in the reference's
    loop order, so the result is the reference's bit for bit given the same tile
    outputs.  Returns (probability float64, map bool) on pred's device; voxels no
    tile covers get 0."""
    if pred.device.type != "cuda":
        raise ValueError("the tile vote runs on the device: pass the tile outputs as a HIP tensor")
    l, h, w = shape
    dev = pred.device
    ny = math.ceil(h / (HEIGHT // STEP_RATIO))
    nx = math.ceil(w / (WIDTH // STEP_RATIO))
    ctx = _vote_context(dev)
    prob = torch.empty(shape, dtype=torch.float64, device=dev)
    mask = torch.empty(shape, dtype=torch.uint8, device=dev)
    starts = ([s[0] for s in slices[::ny * nx]], [s[2] for s in slices[:ny * nx:nx]], [s[4] for s in slices[:nx]])
    ctx.tile_vote(pred.contiguous(), (NUM_PICS, HEIGHT, WIDTH), starts, shape, prob, mask)
    return prob, mask.bool()


_contexts: dict = {}


def _vote_context(dev):
    from . import Context

    idx = dev.index if dev.index is not None else 0
    if idx not in _contexts:
        _contexts[idx] = Context(idx)
    return _contexts[idx]


def segment(model, volume: np.ndarray, mode: str | None = None, device="cuda", batch: int = 64):
    """Normalize, run the tiles through the model on the HIP `device`, vote there.
    Returns the float64 probability and the bool map as device tensors."""
    if torch.device(device).type != "cuda":
        raise ValueError("segment runs on a HIP device")
    x = normalize_device(volume, mode, device)
    sl = tile_slices(volume.shape)
    pred = predict_tiles(model, x, sl, batch)
    return vote(pred, volume.shape, sl)


# ------------------------------------------------------ hand-off to the RSF
def seed_volume(binary: torch.Tensor) -> torch.Tensor:
    """The 0/1 map as a 0/255 uint8 seed for phi0 = sdf(seed - 13)."""
    return binary.to(torch.uint8) * 255


def phi0_from_probability(prob: torch.Tensor, scale: float = 1.0, eps: float = 1e-6) -> torch.Tensor:
    """Differentiable phi0 = -scale * logit(p): negative where p > 1/2, the inside
    convention of autograd.bce_inside_loss."""
    p = prob.clamp(eps, 1.0 - eps)
    return -scale * torch.log(p / (1.0 - p))
