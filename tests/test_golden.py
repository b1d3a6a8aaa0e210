"""Golden vectors (tests/golden/rsf_small.npz, written by make_golden.py).

CPU checks: the oracle reproduces every stored vector bit for bit, the stored
taps and Heaviside / Dirac samples agree with their closed forms in FP64, and
the stored one-step results agree with the independent FP64 numpy restatement
of the iteration (tests/np_reference.py)."""
import json
import os

import numpy as np
import pytest

import np_reference as npr
import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rsf_small.npz")


def load():
    z = np.load(GOLDEN)  # allow_pickle stays False: plain arrays only
    return z, json.loads(str(z["manifest"]))


def cases():
    _, man = load()
    return [(c["name"], T) for c in man for T in c["steps"]]


@pytest.mark.parametrize("name,T", cases())
def test_oracle_reproduces_golden(name, T):
    z, man = load()
    c = next(c for c in man if c["name"] == name)
    img, phi0 = z[c["scene"] + "_image"], z[c["scene"] + "_phi0"]
    got = oracle.run(phi0, img.astype(np.float32), oracle.make_params(**c["params"]), T)
    assert np.array_equal(got, z[f"{name}_T{T}"])


@pytest.mark.parametrize("sigma", [1, 3, 5])
def test_golden_taps_closed_form(sigma):
    z, _ = load()
    g64, P, m = npr.taps64(float(sigma))
    g = z[f"taps_s{sigma}"]
    assert g.shape == (P,)
    np.testing.assert_allclose(g, g64, rtol=3e-7)
    assert np.array_equal(g, g[::-1])  # symmetric about miu = P/2


@pytest.mark.parametrize("eps", [0.2, 1.0])
def test_golden_heaviside_dirac_closed_form(eps):
    z, _ = load()
    tag = str(eps).replace(".", "")
    phi = z["phi_samples"].astype(np.float64)
    e = np.float64(np.float32(eps))
    H = 0.5 * (1.0 + (2.0 / np.pi) * np.arctan(phi / e))
    D = (1.0 / np.pi) * (e / (e * e + phi * phi))
    np.testing.assert_allclose(z[f"heaviside_eps{tag}"], H, rtol=0, atol=2e-7)
    np.testing.assert_allclose(z[f"dirac_eps{tag}"], D, rtol=2e-6)


@pytest.mark.parametrize("name", [c for c, T in cases() if T == 1])
def test_golden_one_step_matches_numpy_fp64(name):
    z, man = load()
    c = next(c for c in man if c["name"] == name)
    kw = dict(c["params"])
    img, phi0 = z[c["scene"] + "_image"], z[c["scene"] + "_phi0"]
    g32, _, _ = oracle.taps(kw["sigma"])
    ref = npr.step(phi0, img, sigma=kw["sigma"], lambda1=kw["lambda1"], lambda2=kw.get("lambda2"), nu=kw["nu"],
                   mu=kw["mu"], dt=kw["dt"], eps=kw["epsilon"], sigma2=kw.get("sigma2", 0.0),
                   g=g32.astype(np.float64))
    got = z[f"{name}_T1"]
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()
