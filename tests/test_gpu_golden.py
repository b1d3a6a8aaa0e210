"""Device engine against the stored golden vectors (tests/golden/rsf_small.npz):
the literal mode reproduces them bit for bit; the fast mode meets the one-step
gate max|dphi| <= 1e-4 * max|phi| and, after T > 1 steps, identical sign masks
with values within 1e-3 * max|phi|."""
import numpy as np
import pytest

import rsf_amd
from gpu_util import one_step_gate, to_dev
from test_golden import cases, load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import torch  # noqa: F401

    c = rsf_amd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("mode", [rsf_amd.MODE_FAST, rsf_amd.MODE_LITERAL])
@pytest.mark.parametrize("name,T", cases())
def test_engine_matches_golden(ctx, name, T, mode):
    import torch

    z, man = load()
    c = next(c for c in man if c["name"] == name)
    img, phi0 = z[c["scene"] + "_image"], z[c["scene"] + "_phi0"]
    ref = z[f"{name}_T{T}"]
    p = rsf_amd.make_params(mode=mode, iterations=T, **c["params"])
    print(f"golden {name} T={T} mode={mode} shape={phi0.shape}", flush=True)
    d_img, d_phi = to_dev(img), to_dev(phi0)
    if T == 1:
        out = torch.empty_like(d_phi)
        ctx.step(d_img, d_phi, out, p)
    else:
        ctx.run(d_img, d_phi, p)
        out = d_phi
    got = out.cpu().numpy()
    scale = float(np.abs(ref).max())
    err = float(np.abs(got.astype(np.float64) - ref).max())
    if mode == rsf_amd.MODE_LITERAL:
        assert np.array_equal(got, ref), err / scale      # bit-identical arithmetic
    elif T == 1:
        e, tol = one_step_gate(got, ref)
        assert e <= tol, (e, tol)
    else:
        assert err <= 1e-3 * scale, err / scale
        assert int(np.count_nonzero((got > 0) != (ref > 0))) == 0
