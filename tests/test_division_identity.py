"""The literal Heaviside takes phi / eps as phi * (1/eps) when eps is a power of
two whose reciprocal is a normal float (the kernels' heav_literal_r, with the
reciprocal from eps_reciprocal in csrc/rsf_sweep.hip).  Both operations round the
exact value phi * 2^-k once, so they must agree on every float: checked here
with IEEE float32 arithmetic over normal, subnormal, huge and special inputs,
and for the powers of two at the ends of the accepted range."""
import numpy as np


def _samples():
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**32, size=2_000_000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    x = x[np.isfinite(x)]
    specials = np.array([0.0, -0.0, np.inf, -np.inf, 1.0, -1.0, 3.0, 1e-45, -1e-45, 1.17549435e-38,
                         3.4028235e38, -3.4028235e38, 2.0**-149, 2.0**-126, 2.0**127], np.float32)
    return np.concatenate([x, specials])


def test_division_by_a_power_of_two_equals_multiplication_by_its_reciprocal():
    x = _samples()
    # eps = 2^k for the k the host accepts (eps and 1/eps normal): k in [-126, 126]
    for k in list(range(-126, -110)) + list(range(-8, 9)) + list(range(110, 127)):
        eps = np.float32(2.0**k)
        r = np.float32(1.0) / eps
        assert np.isfinite(r) and r == np.float32(2.0**-k)
        with np.errstate(over="ignore", under="ignore"):
            q_div = x / eps
            q_mul = x * r
        assert np.array_equal(q_div.view(np.uint32), q_mul.view(np.uint32)), k


def test_other_eps_are_not_exact_reciprocals():
    """0.2 (the reference's value) and 3 have no exact float reciprocal: the
    multiplication would differ somewhere, which is why they keep the division."""
    x = _samples()
    for eps in (np.float32(0.2), np.float32(3.0)):
        with np.errstate(over="ignore", under="ignore"):
            differ = (x / eps).view(np.uint32) != (x * (np.float32(1.0) / eps)).view(np.uint32)
        assert differ.any()
