"""The FP64 adjoint chain at the differentiated workload's real size, against
FP64 central differences stored in tests/golden/config4_fd.npz (written by
tests/golden/make_config4_fd.py; about an hour of CPU, so the differences are
stored and the adjoint is recomputed here).

Volume 0 of the batch (128^3, 40 tubes, seed 4000), sigma 3, lambda 1,
nu 65.025, mu 1, dt 0.1, eps 1, loss = mean BCE(sigmoid(-phi_K/eps), mask):
* phi0 gradient after K = 32 steps at 64 voxels: where the differences at
  h = 1e-3 and 1e-4 agree (converged), the adjoint equals them within 1e-3;
* parameter gradient after K = 8 steps: the differences at relative steps 1e-6
  and 1e-7 converge to the adjoint.  (After 32 steps the loss is not smooth in
  the parameters at such scales: the differences do not converge there.)
The GPU gradient is checked against this FP64 adjoint chain in
tests/test_gpu_workloads.py."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_config4_fd as m  # noqa: E402

FIX = os.path.join(HERE, "golden", "config4_fd.npz")


def test_fp64_adjoint_matches_central_differences_at_config4_size():
    d = np.load(FIX)
    img, phi0, mask = m.scene()
    p = m.params(m.THETA)
    states = m.forward(phi0, img, p)
    assert abs(m.loss(states[-1], mask) - float(d["loss"])) <= 1e-12 * float(d["loss"])
    g, _ = m.adjoint(states, img, mask, p)
    g = g.ravel()[d["idx"]]
    # the same adjoint as when the differences were taken
    np.testing.assert_allclose(g, d["grad"], rtol=1e-9, atol=1e-9 * float(np.abs(d["grad"]).max()))
    fd = d["fd"]
    scale = np.maximum(np.abs(fd[:, 1]), 1e-3 * float(np.abs(g).max()))
    converged = np.abs(fd[:, 0] - fd[:, 1]) <= 1e-3 * scale
    assert converged.sum() >= 56, int(converged.sum())
    err = np.abs(g - fd[:, 0]) / scale
    assert float(err[converged].max()) <= 1e-3, float(err[converged].max())
    # K = 8: parameter gradient
    st8 = m.forward(phi0, img, p, int(d["theta_k"]))
    _, gth = m.adjoint(st8, img, mask, p)
    np.testing.assert_allclose(gth, d["theta_grad"], rtol=1e-9)
    tfd = d["theta_fd"]
    assert np.all(np.abs(tfd[:, 0] - tfd[:, 1]) <= 1e-2 * np.abs(tfd[:, 1]))   # converging
    np.testing.assert_allclose(gth, tfd[:, 1], rtol=1e-4)
