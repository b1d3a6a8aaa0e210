import os
import sys

import numpy as np
import pytest

try:  # load torch's HIP runtime first so librsf binds to the same one
    import torch  # noqa: F401
except Exception:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "paper-rsf-vesselseg_amd")
for p in (ROOT, PKG, os.path.dirname(os.path.abspath(__file__))):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def small_scene(shape=(24, 26, 28), seed=0, tubes=3, noise=10.0):
    """Small (Z,Y,X) test scene: a few straight tubes at 200 on 30 plus noise,
    and phi0 = signed distance (negative inside) to the thresholded tubes."""
    from scipy import ndimage

    rng = np.random.default_rng(seed)
    l, h, w = shape
    z, y, x = np.meshgrid(np.arange(l), np.arange(h), np.arange(w), indexing="ij")
    img = np.full(shape, 30.0)
    for _ in range(tubes):
        p0 = rng.uniform([0, 0, 0], [l, h, w])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        r = rng.uniform(2.0, 4.0)
        v = np.stack([z - p0[0], y - p0[1], x - p0[2]], -1)
        t = v @ d
        dist = np.linalg.norm(v - t[..., None] * d, axis=-1)
        img[dist <= r] = 200.0
    img += np.linspace(-20, 20, w)[None, None, :]
    img += rng.normal(0, noise, shape)
    img = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    mask = ndimage.gaussian_filter(img.astype(np.float64), 1.0) > 110
    if mask.all() or not mask.any():
        mask = np.zeros(shape, bool)
        mask[l // 2 - 3: l // 2 + 3, h // 2 - 3: h // 2 + 3, :] = True
    din = ndimage.distance_transform_edt(mask)
    dout = ndimage.distance_transform_edt(~mask)
    phi = np.where(mask, -(din - 0.5), dout - 0.5).astype(np.float32)
    return img, phi


@pytest.fixture
def scene():
    return small_scene()
