"""Block grid and overlap blend, numpy -- TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference's block workflow (README.md:6-14,56-59): a
volume is cut into overlapping blocks, each block is evolved on its own, and
the blocks are blended back with linear ramps over the overlaps
(interpolating_volumes.py:10-40, which merges two halves along y).  The
reference's disassemble.py / reassemble.py are not in the repository
(README.md:56,59 name them); the grid below is this library's own layout, and
the blend generalises interpolate_volumes to every axis as a tensor product of
its ramps.  Only the test suite imports this module.

Parity pins (tests/test_blocks.py):
* blend of two blocks along y == interpolate_volumes(A, B, k) of
  interpolating_volumes.py:10-40 (restated there verbatim in behaviour), in
  float64, bit for bit;
* the tensor-product blend agrees with blending along x, then y, then z with
  the two-block formula (up to float64 rounding);
* disassemble -> reassemble is the identity, and a constant volume blends to
  itself;
* the grid covers every voxel by one or two blocks per axis.

Arrays are (l, h, w) = (Z, Y, X) numpy volumes.
"""
from __future__ import annotations

import numpy as np


def axis_grid(n: int, B: int, k: int, min_size: int):
    """[(start, size)] of one axis; see block_grid."""
    if B >= n:
        return [(0, n)]
    if B < 2 * k + 1:
        raise ValueError("block size must exceed twice the overlap")
    S = B - k
    g, s = [], 0
    while True:
        if s + B >= n:
            g.append((s, n - s))
            break
        g.append((s, B))
        s += S
    if len(g) >= 2 and g[-1][1] < max(min_size, 2 * k):
        g.pop()
        g[-1] = (g[-1][0], n - g[-1][0])
    return g


def block_grid(shape, block_size, overlap, min_size=1):
    """Blocks of a (l, h, w) volume with block_size = (bw, bh, bl).

    Along each axis: starts j*(B - k) with size B, the last block clipped to
    the volume, a last block shorter than max(min_size, 2k) folded into the
    previous one; B >= n gives one block.  Returns (blocks, origins, axes):
    blocks [(offset, w, h, l)] packed x-fastest ((jz*Jy+jy)*Jx+jx), origins
    (n, 3) int64 of (x0, y0, z0), axes the three per-axis grids (x, y, z)."""
    l, h, w = shape
    axes = [axis_grid(n, B, overlap, min_size) for n, B in zip((w, h, l), block_size)]
    blocks, origins, off = [], [], 0
    for z0, bl in axes[2]:
        for y0, bh in axes[1]:
            for x0, bw in axes[0]:
                blocks.append((off, bw, bh, bl))
                origins.append((x0, y0, z0))
                off += bw * bh * bl
    return blocks, np.array(origins, np.int64).reshape(-1, 3), axes


def disassemble(volume, blocks, origins):
    """Packed copy of the blocks of `volume` (same dtype)."""
    total = max(o + w * h * l for (o, w, h, l) in blocks)
    out = np.zeros(total, volume.dtype)
    for (off, w, h, l), (x0, y0, z0) in zip(blocks, origins):
        out[off:off + w * h * l] = volume[z0:z0 + l, y0:y0 + h, x0:x0 + w].ravel()
    return out


def axis_weights(grid, j, k):
    """Weights of block j along one axis, per coordinate of the block: 1, with
    the ramp numpy.linspace(0, 1, k) over the first k coordinates (rising)
    when a block precedes it and 1 - ramp over the last k (falling) when one
    follows it -- the two factors of interpolating_volumes.py:32-35."""
    size = grid[j][1]
    wgt = np.ones(size, np.float64)
    if k > 0:
        alpha = np.linspace(0, 1, k)
        if j > 0:
            wgt[:k] = alpha
        if j + 1 < len(grid):
            # overlap with block j+1: the coordinates [start_{j+1}, start_{j+1}+k) of
            # the volume, i.e. positions start_{j+1}-start_j .. +k of this block
            p = grid[j + 1][0] - grid[j][0]
            wgt[p:p + k] = 1.0 - alpha
    return wgt


def reassemble(packed, shape, blocks, origins, axes, overlap):
    """Blend of the packed float32 blocks into a float32 (l, h, w) volume:
    sum over the blocks covering a voxel of ((wz*wy)*wx) * phi, in float64,
    blocks taken in packed order, rounded once to float32."""
    l, h, w = shape
    acc = np.zeros(shape, np.float64)
    Jx, Jy = len(axes[0]), len(axes[1])
    for bi, ((off, bw, bh, bl), (x0, y0, z0)) in enumerate(zip(blocks, origins)):
        jx, jy, jz = bi % Jx, (bi // Jx) % Jy, bi // (Jx * Jy)
        wx = axis_weights(axes[0], jx, overlap)
        wy = axis_weights(axes[1], jy, overlap)
        wz = axis_weights(axes[2], jz, overlap)
        wgt = (wz[:, None, None] * wy[None, :, None]) * wx[None, None, :]
        f = packed[off:off + bw * bh * bl].reshape(bl, bh, bw).astype(np.float64)
        acc[z0:z0 + bl, y0:y0 + bh, x0:x0 + bw] += wgt * f
    return acc.astype(np.float32)
