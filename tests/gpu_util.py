"""Helpers for the GPU tests: device tensors, the oracle gate, scenes."""
import numpy as np

import oracle
import rsf_amd


def dev():
    import torch

    return torch.device("cuda:0")


def to_dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a)).to(dev())


def scene(shape, tubes=8, seed=1, rmin=2.0, rmax=6.0, noise=15.0):
    img, mask = rsf_amd.synth_tubes(shape, tubes=tubes, rmin=rmin, rmax=rmax, noise=noise, seed=seed)
    return img, rsf_amd.initial_phi(img), mask


def oracle_params(p: rsf_amd.Params):
    return oracle.make_params(sigma=p.sigma, lambda1=p.lambda1, lambda2=p.lambda2, nu=p.nu, mu=p.mu,
                              dt=p.dt, epsilon=p.epsilon, sigma2=p.sigma2)


def one_step_gate(got, ref, rel=1e-4):
    """The one-step parity gate: max|dphi| <= rel * max|phi|."""
    err = float(np.abs(got.astype(np.float64) - ref).max())
    tol = rel * float(np.abs(ref).max())
    return err, tol


def mask_disagreement(a, b):
    return float(np.mean((a > 0) != (b > 0)))
