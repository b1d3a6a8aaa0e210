"""The rsf command line on the GPU: the plain run, the --block pipeline and the
--batch manifest equal the library calls they stand for, and literal mode
equals the oracle.  The command line runs in this process (rsf_cli_main, the
code of the rsf executable): the test process already uses the GPU and starts
no child."""
import numpy as np
import pytest

import oracle
import rsf_amd
from gpu_util import oracle_params, run_cli, scene, to_dev
from oracle import blocks as ob

pytestmark = pytest.mark.gpu

ARGS = ["--sigma", "1", "--t", "3", "--dt", "0.1", "--meu", "1", "--e", "1", "--v", "65.025", "--epsilon", "1",
        "--mode", "literal"]


def _cli(*args, cwd=None):
    return run_cli(*args, cwd=cwd)


def _params():
    return rsf_amd.make_params(sigma=1.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0,
                               mode=rsf_amd.MODE_LITERAL, iterations=3)


def _seed(shape, seed):
    img, phi0, mask = scene(shape, tubes=5, seed=seed)
    # a seed volume whose sdf(seed - 13) is the starting level set (main.cpp:174-183)
    return img, (mask.astype(np.uint8) * 255)


def test_cli_plain_run_equals_oracle(tmp_path):
    shape = (26, 30, 34)
    img, seed = _seed(shape, 31)
    np.save(tmp_path / "img.npy", img)
    np.save(tmp_path / "seed.npy", seed)
    r = _cli("--image", "img.npy", "--binary", "seed.npy", "--ouput", "phi.npy", *ARGS, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("It takes") == 3
    got = np.load(tmp_path / "phi.npy")
    phi0 = rsf_amd.sdf_from_seed(seed.astype(np.float32), 13.0)
    ref = oracle.run(phi0, img.astype(np.float32), oracle_params(_params()), 3)
    assert got.dtype == np.float32 and np.array_equal(got, ref)


def test_cli_block_mode_equals_library_pipeline(tmp_path):
    import torch

    shape = (40, 44, 50)
    img, seed = _seed(shape, 32)
    np.save(tmp_path / "img.npy", img)
    np.save(tmp_path / "seed.npy", seed)
    r = _cli("--image", "img.npy", "--binary", "seed.npy", "--ouput", "phi.npy", "--block", "30,26,24",
             "--overlap", "6", *ARGS, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    phi0 = rsf_amd.sdf_from_seed(seed.astype(np.float32), 13.0)
    ctx = rsf_amd.Context(0)
    d = to_dev(phi0)
    ctx.run_volume_blocks(to_dev(img), d, (30, 26, 24), 6, _params())
    torch.cuda.synchronize()
    ctx.close()
    assert np.array_equal(np.load(tmp_path / "phi.npy"), d.cpu().numpy())


def test_cli_batch_manifest_equals_oracle_per_block(tmp_path):
    shapes = [(24, 30, 33), (30, 24, 27), (26, 26, 40)]
    lines, refs = [], []
    for k, shp in enumerate(shapes):
        img, seed = _seed(shp, 40 + k)
        (tmp_path / "blocks").mkdir(exist_ok=True)
        np.save(tmp_path / "blocks" / f"img{k}.npy", img if k != 1 else img.astype(np.float32))
        np.save(tmp_path / "blocks" / f"seed{k}.npy", seed)
        lines.append(f"img{k}.npy seed{k}.npy phi{k}.npy")
        phi0 = rsf_amd.sdf_from_seed(seed.astype(np.float32), 13.0)
        refs.append(oracle.run(phi0, img.astype(np.float32), oracle_params(_params()), 3))
    (tmp_path / "blocks" / "manifest.txt").write_text("# image seed output\n" + "\n".join(lines) + "\n")
    r = _cli("--batch", "blocks/manifest.txt", *ARGS, cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("It takes") == 3
    for k, ref in enumerate(refs):
        assert np.array_equal(np.load(tmp_path / "blocks" / f"phi{k}.npy"), ref), k


def test_block_files_round_trip_equals_oracle(tmp_path):
    """disassemble_npy -> rsf --batch -> reassemble_npy (the reference's
    disassemble / rsfbatch / reassemble workflow, README.md:56-59) equals the oracle
    evolving every block from its own seed, blended by the oracle's blend."""
    from rsf_amd import blockfiles

    shape, bs, k = (40, 44, 50), (30, 26, 24), 6
    img, seed = _seed(shape, 33)
    np.save(tmp_path / "vol.npy", img)
    np.save(tmp_path / "seed.npy", seed)
    d = tmp_path / "blocks"
    ctx = rsf_amd.Context(0)
    desc = blockfiles.disassemble_npy(str(tmp_path / "vol.npy"), str(d), bs, k, str(tmp_path / "seed.npy"),
                                      min_size=7, context=ctx)
    r = _cli("--batch", str(d / "manifest.txt"), *ARGS)
    assert r.returncode == 0, r.stderr
    blockfiles.reassemble_npy(str(d), str(tmp_path / "phi.npy"), context=ctx)
    ctx.close()
    got = np.load(tmp_path / "phi.npy")
    blocks, origins, axes = ob.block_grid(shape, bs, k, 7)
    assert len(desc["blocks"]) == len(blocks)
    packed = np.empty(sum(b[1] * b[2] * b[3] for b in blocks), np.float32)
    for (off, w, h, l), (x0, y0, z0) in zip(blocks, origins):
        ib = img[z0:z0 + l, y0:y0 + h, x0:x0 + w]
        sb = seed[z0:z0 + l, y0:y0 + h, x0:x0 + w].astype(np.float32)
        phi0 = rsf_amd.sdf_from_seed(np.ascontiguousarray(sb), 13.0)
        packed[off:off + w * h * l] = oracle.run(phi0, np.ascontiguousarray(ib).astype(np.float32),
                                                 oracle_params(_params()), 3).ravel()
    assert np.array_equal(got, ob.reassemble(packed, shape, blocks, origins, axes, k))
