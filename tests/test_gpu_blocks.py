"""GPU parity of the block pipeline (README.md:56-59): disassembly and
reassembly against the oracle bit for bit, and the whole pipeline in literal
mode against the oracle's per-block evolution followed by the oracle's blend."""
import numpy as np
import pytest

import oracle
import rsf_amd
from gpu_util import oracle_params, scene, to_dev
from oracle import blocks as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch  # noqa: F401

    c = rsf_amd.Context(0)
    yield c
    c.close()


GRIDS = [
    ((40, 50, 60), (64, 64, 64), 8, 1),
    ((40, 50, 60), (32, 32, 32), 4, 21),
    ((33, 47, 71), (20, 17, 25), 3, 7),
    ((33, 47, 71), (9, 9, 9), 4, 1),
    ((10, 11, 12), (5, 4, 3), 1, 1),
    ((10, 11, 12), (5, 4, 3), 0, 1),
    ((64, 96, 128), (48, 40, 30), 12, 21),
    ((20, 150, 30), (30, 100, 20), 46, 1),   # the reference's overlap of 46 along y
]


@pytest.mark.parametrize("shape,bs,k,mn", GRIDS)
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_disassemble_matches_oracle(ctx, shape, bs, k, mn, dtype):
    import torch

    rng = np.random.default_rng(11)
    vol = (rng.integers(0, 256, shape) if dtype == np.uint8 else rng.normal(size=shape)).astype(dtype)
    g = rsf_amd.BlockGrid(shape, bs, k, mn)
    ref = ob.disassemble(vol, g.blocks, g.origins)
    d = torch.zeros(g.packed_voxels, dtype=torch.uint8 if dtype == np.uint8 else torch.float32, device="cuda")
    ctx.disassemble(to_dev(vol), g, d)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), ref)


@pytest.mark.parametrize("shape,bs,k,mn", GRIDS)
def test_reassemble_matches_oracle(ctx, shape, bs, k, mn):
    import torch

    rng = np.random.default_rng(12)
    g = rsf_amd.BlockGrid(shape, bs, k, mn)
    _, _, axes = ob.block_grid(shape, bs, k, mn)
    packed = (rng.normal(size=g.packed_voxels) * 50).astype(np.float32)
    ref = ob.reassemble(packed, shape, g.blocks, g.origins, axes, k)
    out = torch.full(shape, float("nan"), dtype=torch.float32, device="cuda")
    ctx.reassemble(to_dev(packed), g, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), ref)


def test_disassemble_accepts_any_block_list(ctx):
    import torch

    rng = np.random.default_rng(13)
    shape = (30, 31, 32)
    vol = rng.normal(size=shape).astype(np.float32)
    blocks, origins, off = [], [], 0
    for _ in range(9):
        dims = rng.integers(1, 12, 3)
        o = [int(rng.integers(0, shape[2 - a] - dims[a] + 1)) for a in range(3)]
        blocks.append((off, *map(int, dims)))
        origins.append(o)
        off += int(np.prod(dims))
    g = type("G", (), {})()
    g.blocks, g.c_blocks = blocks, rsf_amd.blocks_array(blocks)
    g.c_origins = np.array(origins, np.int64)
    d = torch.zeros(off, dtype=torch.float32, device="cuda")
    ctx.disassemble(to_dev(vol), g, d)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy(), ob.disassemble(vol, blocks, g.c_origins))


def test_block_tables_are_checked(ctx):
    import torch

    vol = torch.zeros((10, 10, 10), device="cuda")
    g = type("G", (), {})()
    g.blocks = [(0, 5, 5, 5)]
    g.c_blocks = rsf_amd.blocks_array(g.blocks)
    g.c_origins = np.array([[6, 0, 0]], np.int64)   # x0 + w > 10
    with pytest.raises(rsf_amd.RSFError):
        ctx.disassemble(vol, g, torch.zeros(125, device="cuda"))
    g.blocks = [(8, 2, 2, 2), (0, 2, 2, 2)]                # offsets must ascend
    g.c_blocks = rsf_amd.blocks_array(g.blocks)
    g.c_origins = np.zeros((2, 3), np.int64)
    with pytest.raises(rsf_amd.RSFError):
        ctx.disassemble(vol, g, torch.zeros(16, device="cuda"))
    g.blocks = [(0, 2, 2, 2), (7, 2, 2, 2)]                # and the blocks disjoint
    g.c_blocks = rsf_amd.blocks_array(g.blocks)
    with pytest.raises(rsf_amd.RSFError):
        ctx.disassemble(vol, g, torch.zeros(16, device="cuda"))


@pytest.mark.parametrize("shape,bs,k", [((40, 44, 50), (30, 26, 24), 6), ((28, 60, 31), (40, 36, 40), 12)])
def test_volume_pipeline_matches_oracle_blocks(ctx, shape, bs, k):
    """Literal mode: every block evolves exactly as the oracle evolves it alone,
    and the blend is the oracle's -- the pipeline equals the oracle's one."""
    import torch

    img, phi0, _ = scene(shape, tubes=6, seed=21)
    p = rsf_amd.make_params(sigma=1.0, epsilon=1.0, mode=rsf_amd.MODE_LITERAL, iterations=4)
    P, _ = rsf_amd.taps(1.0)[1:]
    g = rsf_amd.BlockGrid(shape, bs, k, P)
    _, _, axes = ob.block_grid(shape, bs, k, P)
    assert len(g.blocks) >= 2
    pi = ob.disassemble(img, g.blocks, g.origins)
    pp = ob.disassemble(phi0, g.blocks, g.origins)
    for (off, w, h, l) in g.blocks:
        n = w * h * l
        pp[off:off + n] = oracle.run(pp[off:off + n].reshape(l, h, w),
                                     pi[off:off + n].reshape(l, h, w).astype(np.float32),
                                     oracle_params(p), 4).ravel()
    ref = ob.reassemble(pp, shape, g.blocks, g.origins, axes, k)
    d_phi = to_dev(phi0)
    ctx.run_volume_blocks(to_dev(img), d_phi, bs, k, p)
    torch.cuda.synchronize()
    assert np.array_equal(d_phi.cpu().numpy(), ref)


def test_volume_pipeline_single_block_equals_plain_run(ctx):
    import torch

    shape = (36, 40, 44)
    img, phi0, _ = scene(shape, tubes=5, seed=22)
    p = rsf_amd.make_params(sigma=3.0, epsilon=1.0, iterations=5)
    a, b = to_dev(phi0), to_dev(phi0)
    d_img = to_dev(img)
    ctx.run_volume_blocks(d_img, a, (64, 64, 64), 8, p)
    ctx.run(d_img, b, p)
    torch.cuda.synchronize()
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
