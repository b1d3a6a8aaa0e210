"""Helpers for the GPU tests: device tensors, the oracle gate, scenes."""
import os

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


class CliResult:
    def __init__(self, returncode, stdout, stderr):
        self.returncode, self.stdout, self.stderr = returncode, stdout, stderr


_cli_lib = None


def run_cli(*args, cwd=None):
    """The rsf command line run inside this process: rsf_cli_main of librsf_cli.so,
    the code of the rsf executable, so a test process that already uses the GPU
    starts no child.  stdout and stderr are captured at the file-descriptor level."""
    import ctypes
    import sys
    import tempfile

    import rsf_amd

    global _cli_lib
    if _cli_lib is None:
        rsf_amd.lib()   # librsf first, from the same path
        _cli_lib = ctypes.CDLL(os.path.join(rsf_amd.PKG_DIR, "librsf_cli.so"))
        _cli_lib.rsf_cli_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
        _cli_lib.rsf_cli_main.restype = ctypes.c_int
    argv = [b"rsf"] + [os.fsencode(str(a)) for a in args]
    arr = (ctypes.c_char_p * (len(argv) + 1))(*argv, None)
    here = os.getcwd()
    with tempfile.TemporaryFile() as out, tempfile.TemporaryFile() as err:
        sys.stdout.flush()
        sys.stderr.flush()
        keep1, keep2 = os.dup(1), os.dup(2)
        os.dup2(out.fileno(), 1)
        os.dup2(err.fileno(), 2)
        try:
            if cwd is not None:
                os.chdir(cwd)
            rc = _cli_lib.rsf_cli_main(len(argv), arr)
        finally:
            os.chdir(here)
            os.dup2(keep1, 1)
            os.dup2(keep2, 2)
            os.close(keep1)
            os.close(keep2)
        out.seek(0)
        err.seek(0)
        return CliResult(rc, out.read().decode(), err.read().decode())
