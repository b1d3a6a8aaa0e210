"""Writes tests/golden/rsf_small.npz: golden vectors of the RSF iteration.

Inputs are small seeded scenes (tests/conftest.py:small_scene); expected
outputs come from the oracle (the CPU restatement of CUDA_RSF/main.cpp:222-300
and CUDA_RSF/kernel.cu), whose functions are pinned by the closed forms, brute
force and the independent FP64 numpy restatement in tests/test_oracle.py.  The
fixtures freeze those outputs so that the oracle cannot drift unnoticed and so
that the device engine can be checked against stored vectors.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import oracle  # noqa: E402
from conftest import small_scene  # noqa: E402

# (name, scene shape (Z,Y,X), scene seed, oracle parameters, step counts)
CASES = [
    ("ref_defaults_s3_eps02", (24, 26, 28), 11,
     dict(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=0.2), (1, 5)),
    ("s1_eps1", (24, 26, 28), 11,
     dict(sigma=1.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0), (1, 5)),
    ("s3_eps1_lambda_sigma2", (24, 26, 28), 11,
     dict(sigma=3.0, lambda1=1.3, lambda2=0.7, nu=6.5, mu=0.6, dt=0.1, epsilon=1.0, sigma2=1.0), (1,)),
    ("s3_eps1_noncubic", (30, 25, 27), 12,
     dict(sigma=3.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0), (1, 3)),
    ("s5_eps1", (40, 38, 37), 13,
     dict(sigma=5.0, lambda1=1.0, nu=65.025, mu=1.0, dt=0.1, epsilon=1.0), (1,)),
]
PHI_SAMPLES = np.array([0.0, 1e-3, -0.05, 0.2, -0.2, 1.0, -3.0, 10.0, -10.0, 123.5], np.float32)


def main():
    oracle.build()
    out = {}
    manifest = []
    for sigma in (1.0, 3.0, 5.0):
        g, P, m = oracle.taps(sigma)
        out[f"taps_s{int(sigma)}"] = g
    out["phi_samples"] = PHI_SAMPLES
    for eps in (0.2, 1.0):
        tag = str(eps).replace(".", "")
        out[f"heaviside_eps{tag}"] = oracle.heaviside(PHI_SAMPLES, eps)
        out[f"dirac_eps{tag}"] = oracle.deri_heaviside(PHI_SAMPLES, eps)
    scenes = {}
    for name, shape, seed, kw, steps in CASES:
        key = f"scene_{seed}_{'x'.join(map(str, shape))}"
        if key not in scenes:
            img, phi = small_scene(shape, seed=seed)
            scenes[key] = True
            out[key + "_image"] = img
            out[key + "_phi0"] = phi
        img, phi0 = out[key + "_image"], out[key + "_phi0"]
        p = oracle.make_params(**kw)
        for T in steps:
            out[f"{name}_T{T}"] = oracle.run(phi0, img.astype(np.float32), p, T)
        manifest.append({"name": name, "scene": key, "params": kw, "steps": list(steps)})
    out["manifest"] = np.array(json.dumps(manifest))
    path = os.path.join(HERE, "rsf_small.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
