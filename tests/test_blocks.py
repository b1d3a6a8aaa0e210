"""Block pipeline on the CPU: the library's grid against the oracle's, and the
oracle's blend pinned against the reference's two-block formula
(interpolating_volumes.py:10-40), sequential blending and identities."""
import itertools

import numpy as np
import pytest

import rsf_amd
from oracle import blocks as ob


def interpolate_volumes(half_1, half_2, k):
    """The reference's two-half blend along y, interpolating_volumes.py:10-40:
    concat(A[:, :-k], (1-alpha) A[:, -k:] + alpha B[:, :k], B[:, k:]) with
    alpha = numpy.linspace(0, 1, k); float64 result for float32 halves."""
    alpha = np.linspace(0, 1, k)[None, :, None]
    mid = (1 - alpha) * half_1[:, -k:, :] + alpha * half_2[:, :k, :]
    return np.concatenate((half_1[:, :-k, :], mid, half_2[:, k:, :]), axis=1)


def blend_axis(parts, k, axis):
    """Chain interpolate_volumes over consecutive parts along `axis`."""
    out = np.moveaxis(parts[0], axis, 1)
    for p in parts[1:]:
        out = interpolate_volumes(out, np.moveaxis(p, axis, 1), k)
    return np.moveaxis(out, 1, axis)


GRID_CASES = [
    # shape (l, h, w), block size (bw, bh, bl), overlap, min_size
    ((40, 50, 60), (64, 64, 64), 8, 1),      # one block
    ((40, 50, 60), (60, 50, 40), 8, 1),      # block == volume
    ((40, 50, 60), (32, 32, 32), 4, 1),
    ((40, 50, 60), (32, 32, 32), 4, 21),     # folded last blocks
    ((33, 47, 71), (20, 17, 25), 3, 7),
    ((33, 47, 71), (9, 9, 9), 4, 1),         # B = 2k + 1
    ((10, 11, 12), (5, 4, 3), 1, 1),
    ((10, 11, 12), (5, 4, 3), 0, 1),         # no overlap: plain tiling
    ((128, 96, 64), (48, 40, 30), 12, 21),
]


@pytest.mark.parametrize("shape,bs,k,mn", GRID_CASES)
def test_grid_matches_oracle(shape, bs, k, mn):
    g = rsf_amd.BlockGrid(shape, bs, k, mn)
    blocks, origins, axes = ob.block_grid(shape, bs, k, mn)
    assert g.blocks == blocks
    assert np.array_equal(g.origins, origins)
    assert g.packed_voxels == sum(w * h * l for (_, w, h, l) in blocks)


@pytest.mark.parametrize("shape,bs,k,mn", GRID_CASES)
def test_grid_covers_every_voxel_once_or_in_one_overlap(shape, bs, k, mn):
    _, _, axes = ob.block_grid(shape, bs, k, mn)
    for n, grid in zip(shape[::-1], axes):
        cover = np.zeros(n, int)
        for s, size in grid:
            assert size >= 1 and s + size <= n
            cover[s:s + size] += 1
        assert cover.min() == 1 and cover.max() <= 2
        # consecutive blocks overlap by exactly k
        for (s0, n0), (s1, _) in zip(grid, grid[1:]):
            assert s0 + n0 - s1 == k
        if len(grid) > 1:
            assert grid[-1][1] >= max(mn, 2 * k)


def test_grid_rejects_block_not_wider_than_two_overlaps():
    with pytest.raises(rsf_amd.RSFError):
        rsf_amd.BlockGrid((40, 40, 40), (16, 16, 16), 8, 1)
    with pytest.raises(ValueError):
        ob.block_grid((40, 40, 40), (16, 16, 16), 8, 1)


@pytest.mark.parametrize("k", [1, 2, 5, 46])
def test_oracle_blend_equals_reference_two_halves(k):
    # two blocks along y, each with its own contents, as interpolate_volumes merges them
    rng = np.random.default_rng(k)
    l, h, w = 7, 3 * k + 40, 9
    bh = (h + k + 1) // 2 + 1
    blocks, origins, axes = ob.block_grid((l, h, w), (w, bh, l), k)
    assert len(blocks) == 2
    packed = rng.normal(size=sum(b[1] * b[2] * b[3] for b in blocks)).astype(np.float32) * 100
    got = ob.reassemble(packed, (l, h, w), blocks, origins, axes, k)
    halves = [packed[o:o + bw * bh_ * bl].reshape(bl, bh_, bw) for (o, bw, bh_, bl) in blocks]
    ref = interpolate_volumes(halves[0], halves[1], k)
    assert ref.dtype == np.float64 and ref.shape == (l, h, w)
    assert np.array_equal(got, ref.astype(np.float32))


def test_oracle_tensor_blend_equals_sequential_axis_blends():
    rng = np.random.default_rng(3)
    shape, bs, k = (30, 34, 38), (16, 14, 13), 3
    blocks, origins, axes = ob.block_grid(shape, bs, k)
    J = [len(a) for a in axes]
    assert min(J) >= 2
    packed = rng.normal(size=sum(b[1] * b[2] * b[3] for b in blocks)).astype(np.float32)
    got = ob.reassemble(packed, shape, blocks, origins, axes, k)
    parts = [packed[o:o + bw * bh * bl].reshape(bl, bh, bw).astype(np.float64) for (o, bw, bh, bl) in blocks]
    # blend along x within each (jz, jy) row, then along y, then along z
    rows = [[blend_axis(parts[(jz * J[1] + jy) * J[0]:(jz * J[1] + jy + 1) * J[0]], k, 2) for jy in range(J[1])]
            for jz in range(J[2])]
    slabs = [blend_axis(r, k, 1) for r in rows]
    ref = blend_axis(slabs, k, 0)
    np.testing.assert_allclose(got, ref, rtol=3e-7, atol=1e-6)


@pytest.mark.parametrize("shape,bs,k,mn", GRID_CASES)
def test_oracle_disassemble_reassemble_is_identity(shape, bs, k, mn):
    rng = np.random.default_rng(len(shape) + k)
    vol = (rng.normal(size=shape) * 10).astype(np.float32)
    blocks, origins, axes = ob.block_grid(shape, bs, k, mn)
    packed = ob.disassemble(vol, blocks, origins)
    assert np.array_equal(ob.reassemble(packed, shape, blocks, origins, axes, k), vol)
    const = np.full(shape, 0.7, np.float32)
    packed = ob.disassemble(const, blocks, origins)
    assert np.array_equal(ob.reassemble(packed, shape, blocks, origins, axes, k), const)


def test_linspace_restatement_of_the_kernel():
    # the reassembly kernel evaluates alpha_i as i * (1/(k-1)) with alpha_{k-1} = 1
    # (and alpha_0 = 0 for k = 1), which must be numpy.linspace(0, 1, k) exactly
    for k in itertools.chain(range(1, 300), [1000, 4097]):
        step = 1.0 / (k - 1) if k > 1 else 0.0
        mine = np.array([1.0 if (k > 1 and i == k - 1) else i * step for i in range(k)])
        assert np.array_equal(mine, np.linspace(0, 1, k)), k


def test_grid_rejects_absurd_block_counts():
    with pytest.raises(rsf_amd.RSFError, match="2\\^24"):
        rsf_amd.BlockGrid((1024, 1024, 1024), (2, 2, 2), 0)
