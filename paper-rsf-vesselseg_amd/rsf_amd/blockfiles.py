"""On-disk block format of the gigavoxel workflow (README.md:6-14,56-59).

The reference cuts a volume into overlapping `.npy` blocks (disassemble.py),
runs `rsf3d` on each block (rsfbatch.py) and blends the blocks' phi back
(reassemble.py, with the two-block blend of interpolating_volumes.py:10-40);
the three scripts are not in its repository.  Here:

* `disassemble_npy` writes the image (and seed) blocks of the grid
  `rsf_block_grid` describes, `grid.json` (the volume shape, block size,
  overlap and every block's origin) and, with a seed, `manifest.txt` -- one
  `image seed output` line per block, the format of `rsf --batch`;
* `rsf --batch <dir>/manifest.txt ...` evolves all blocks in one batched
  launch and writes each block's phi;
* `reassemble_npy` blends the phi blocks back into one volume with the linear
  ramps of `rsf_reassemble_device`.

Cutting and blending run on the device (`rsf_disassemble_device`,
`rsf_reassemble_device`); the files are read and written through the streaming
`.npy` entry points where whole volumes are involved.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import BlockGrid, Context, npy_info


def _name(kind: str, i: int) -> str:
    return f"{kind}_{i:05d}.npy"


def disassemble_npy(volume_path: str, out_dir: str, block_size, overlap: int, seed_path: str | None = None,
                    min_size: int = 21, context: Context | None = None, device: int = 0) -> dict:
    """Cut `volume_path` (and `seed_path`) into overlapping block files in `out_dir`.

    block_size: (bw, bh, bl) or one int; min_size: the shortest last block along an axis
    (the tap count of the run, 21 at sigma 3).  Returns the grid description written to
    grid.json."""
    import torch

    shape, u8 = npy_info(volume_path)
    if len(shape) != 3:
        raise ValueError(f"{volume_path}: a 3-D volume (Z, Y, X) is required")
    if isinstance(block_size, int):
        block_size = (block_size,) * 3
    ctx = context or Context(device)
    dev = torch.device("cuda", device)
    grid = BlockGrid(shape, block_size, overlap, min_size)
    os.makedirs(out_dir, exist_ok=True)
    sources = [("img", volume_path, u8)]
    if seed_path is not None:
        sshape, su8 = npy_info(seed_path)
        if sshape != shape:
            raise ValueError("seed and image shapes differ")
        sources.append(("seed", seed_path, su8))
    for kind, path, is_u8 in sources:
        dt = torch.uint8 if is_u8 else torch.float32
        vol = torch.empty(shape, dtype=dt, device=dev)
        ctx.npy_load(path, vol)
        packed = torch.empty(grid.packed_voxels, dtype=dt, device=dev)
        ctx.disassemble(vol, grid, packed)
        del vol
        host = packed.cpu().numpy()
        for i, (off, w, h, l) in enumerate(grid.blocks):
            np.save(os.path.join(out_dir, _name(kind, i)), host[off:off + w * h * l].reshape(l, h, w))
    desc = {"shape": list(shape), "block_size": list(block_size), "overlap": int(overlap), "min_size": int(min_size),
            "blocks": [{"origin": [int(v) for v in o], "dims": [int(w), int(h), int(l)]}
                       for o, (_, w, h, l) in zip(grid.origins, grid.blocks)]}
    with open(os.path.join(out_dir, "grid.json"), "w") as f:
        json.dump(desc, f, indent=1)
    if seed_path is not None:   # the batch run needs a seed per block
        with open(os.path.join(out_dir, "manifest.txt"), "w") as f:
            f.write("# image seed output -- rsf --batch manifest.txt\n")
            for i in range(len(grid.blocks)):
                f.write(f"{_name('img', i)} {_name('seed', i)} {_name('phi', i)}\n")
    return desc


def reassemble_npy(block_dir: str, out_path: str, context: Context | None = None, device: int = 0) -> None:
    """Blend the phi blocks of `block_dir` (named as in its manifest, described by its
    grid.json) into one float32 volume at `out_path`."""
    import torch

    with open(os.path.join(block_dir, "grid.json")) as f:
        desc = json.load(f)
    shape = tuple(desc["shape"])
    grid = BlockGrid(shape, tuple(desc["block_size"]), desc["overlap"], desc["min_size"])
    if len(grid.blocks) != len(desc["blocks"]):
        raise ValueError("grid.json does not match the grid its parameters give")
    ctx = context or Context(device)
    dev = torch.device("cuda", device)
    packed = np.empty(grid.packed_voxels, np.float32)
    for i, (off, w, h, l) in enumerate(grid.blocks):
        a = np.load(os.path.join(block_dir, _name("phi", i)))
        if a.shape != (l, h, w):
            raise ValueError(f"{_name('phi', i)}: shape {a.shape}, the grid expects {(l, h, w)}")
        packed[off:off + a.size] = a.astype(np.float32, copy=False).ravel()
    out = torch.empty(shape, dtype=torch.float32, device=dev)
    ctx.reassemble(torch.from_numpy(packed).to(dev), grid, out)
    ctx.npy_save(out_path, out)
