"""The oracle's tile vote of the U-Net seed producer (test infrastructure).

Restates Unet/run_model.py:86-91 (count / total in float64, every voxel's tile
outputs summed in the loop's order, then probability >= 0.5, run_model.py:
29-35) as one exact torch gather per combination of covering tiles, on host
tensors.  tests/test_unet.py pins it against a literal restatement of the
reference loop; the GPU tests check the library's device vote against it."""
from __future__ import annotations

import math

import numpy as np
import torch


def _cover(starts, ends, n):
    """Per coordinate of an axis: the tiles covering it in loop order (index,
    offset), padded with index -1."""
    starts, ends = np.asarray(starts), np.asarray(ends)
    c = np.arange(n)[:, None]
    hit = (c >= starts[None, :]) & (c < ends[None, :])
    m = int(hit.sum(1).max())
    idx = np.full((n, m), -1, np.int64)
    off = np.zeros((n, m), np.int64)
    for q in range(n):
        t = np.nonzero(hit[q])[0]
        idx[q, :len(t)] = t
        off[q, :len(t)] = q - starts[t]
    return torch.from_numpy(idx), torch.from_numpy(off)


def tile_vote(pred: torch.Tensor, shape, slices, tile=(16, 64, 64), stride=(16, 32, 32)):
    """(probability float64, map bool) of the tile outputs pred (T, tz, ty, tx),
    tiles listed z-major as run_model.py:82-85 visits them."""
    tz, ty, tx = tile
    l, h, w = shape
    nz, ny, nx = math.ceil(l / stride[0]), math.ceil(h / stride[1]), math.ceil(w / stride[2])
    pred = pred.detach().cpu()
    zi, zo = _cover([s[0] for s in slices[::ny * nx]], [s[1] for s in slices[::ny * nx]], l)
    yi, yo = _cover([s[2] for s in slices[:ny * nx:nx]], [s[3] for s in slices[:ny * nx:nx]], h)
    xi, xo = _cover([s[4] for s in slices[:nx]], [s[5] for s in slices[:nx]], w)
    count = torch.zeros(shape, dtype=torch.float64)
    total = torch.zeros(shape, dtype=torch.float64)
    flat = pred.reshape(-1)
    for a in range(zi.shape[1]):
        for b in range(yi.shape[1]):
            for c in range(xi.shape[1]):
                ti = (zi[:, a, None, None] * ny + yi[None, :, b, None]) * nx + xi[None, None, :, c]
                ok = (zi[:, a, None, None] >= 0) & (yi[None, :, b, None] >= 0) & (xi[None, None, :, c] >= 0)
                lin = ((ti * tz + zo[:, a, None, None]) * ty + yo[None, :, b, None]) * tx + xo[None, None, :, c]
                v = flat[torch.where(ok, lin, torch.zeros_like(lin))].double()
                count += torch.where(ok, v, torch.zeros_like(v))
                total += ok.double()
    prob = torch.where(total > 0, count / torch.where(total > 0, total, torch.ones_like(total)),
                       torch.zeros_like(count))
    return prob, prob >= 0.5
