"""Writes tests/golden/config4_fd.npz: FP64 central differences of the
differentiated workload, the check of the FP64 adjoint chain at its real size.

Workload: one 128^3 uint8 tube phantom (40 tubes, radius 2-6, +-20 ramp, noise
15, seed 4000 -- volume 0 of the differentiated batch), phi0 the signed distance
of the thresholded sigma-1 blur, sigma 3 (21 taps), lambda1 = lambda2 = 1,
nu = 65.025, mu = 1, dt = 0.1, epsilon = 1, K steps of the reference iteration
(CUDA_RSF/main.cpp:234-300) in FP64 (the oracle's FP64 twin), loss
L = mean over the voxels of BCE(sigmoid(-phi_K / eps), mask).

Part "phi" (K = 32): the 64 voxels (32 uniform, 32 uniform among |phi0| <= 3,
numpy default_rng(4004)), central differences of L at h = 1e-3 and h = 1e-4 for
each, and the FP64 adjoint gradient at those voxels.  About 50 minutes on 8
cores.

Part "theta" (K = 8): dL/d(lambda1, lambda2, nu, mu) by the adjoint and by
central differences at relative steps 1e-6 and 1e-7 (the parameters are
float32: the difference of the stored values is the denominator).  After 32
steps the map theta -> L is not smooth at those scales -- differences at
relative steps 1e-5, 1e-6, 1e-7 give -108.5, -124.1, -67.1 for lambda1 -- so
the parameter gradient is checked where the differences converge.

Run from the repository root:
    python tests/golden/make_config4_fd.py [out.npz] [phi|theta|all]
A part not recomputed keeps the values already in out.npz.
"""
import os
import sys
import time

import numpy as np
from scipy.special import expit

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "paper-rsf-vesselseg_amd")]

import oracle  # noqa: E402
import rsf_amd  # noqa: E402

N, K, EPS = 128, 32, 1.0
K_THETA = 8
THETA = (1.0, 1.0, 65.025, 1.0)


def scene():
    img, mask = rsf_amd.synth_tubes((N, N, N), tubes=40, rmin=2.0, rmax=6.0, bias=20.0, noise=15.0, seed=4000)
    return img, rsf_amd.initial_phi(img, 1.0, 110.0), mask.astype(np.float64)


def params(theta):
    return oracle.make_params(sigma=3.0, lambda1=theta[0], lambda2=theta[1], nu=theta[2], mu=theta[3], dt=0.1,
                              epsilon=EPS)


def forward(phi0, img, p, steps=None):
    st = [np.asarray(phi0, np.float64)]
    for _ in range(K if steps is None else steps):
        st.append(oracle.step_f64(st[-1], img, p))
    return st


def loss(phiK, mask):
    z = -phiK / EPS
    return float(np.mean(np.logaddexp(0.0, z) - mask * z))


def adjoint(states, img, mask, p):
    """dL/dphi0 and dL/d(lambda1, lambda2, nu, mu) by the FP64 adjoint of each step."""
    zK = -states[-1] / EPS
    u = (expit(zK) - mask) * (-1.0 / EPS) / mask.size
    th = np.zeros(4)
    for t in reversed(range(len(states) - 1)):
        u, gp = oracle.step_adjoint_f64(states[t], img, p, u)
        th += gp
    return u, th


def phi_part(img, phi0, mask):
    t0 = time.time()
    p = params(THETA)
    states = forward(phi0, img, p)
    L0 = loss(states[-1], mask)
    g, _ = adjoint(states, img, mask, p)
    print(f"forward + adjoint {time.time() - t0:.0f} s, loss {L0:.6f}", flush=True)
    rng = np.random.default_rng(4004)
    near = np.flatnonzero(np.abs(phi0.ravel()) <= 3.0)
    idx = np.concatenate([rng.choice(phi0.size, 32, replace=False), rng.choice(near, 32, replace=False)])
    hs = (1e-3, 1e-4)
    fd = np.zeros((len(idx), len(hs)))
    for k, v in enumerate(idx):
        for j, h in enumerate(hs):
            vals = []
            for sgn in (1.0, -1.0):
                q = phi0.astype(np.float64).ravel().copy()
                q[v] += sgn * h
                vals.append(loss(forward(q.reshape(phi0.shape), img, p)[-1], mask))
            fd[k, j] = (vals[0] - vals[1]) / (2 * h)
        print(f"voxel {k + 1}/64 ({int(v)}): adjoint {g.ravel()[v]:.6e} fd {fd[k, 0]:.6e} {fd[k, 1]:.6e} "
              f"[{time.time() - t0:.0f} s]", flush=True)
    return dict(idx=idx, fd=fd, hs=np.array(hs), grad=g.ravel()[idx], loss=L0)


def theta_part(img, phi0, mask):
    p = params(THETA)
    states = forward(phi0, img, p, K_THETA)
    _, gth = adjoint(states, img, mask, p)
    rels = (1e-6, 1e-7)
    thfd = np.zeros((4, len(rels)))
    for i in range(4):
        for j, r in enumerate(rels):
            vals, stored = [], []
            for sgn in (1.0, -1.0):
                th = list(THETA)
                th[i] = float(np.float32(th[i] * (1.0 + sgn * r)))
                stored.append(th[i])
                vals.append(loss(forward(phi0, img, params(th), K_THETA)[-1], mask))
            thfd[i, j] = (vals[0] - vals[1]) / (stored[0] - stored[1])
        print(f"theta {i}: adjoint {gth[i]:.6e} fd {thfd[i]}", flush=True)
    return dict(theta_k=K_THETA, theta_grad=gth, theta_fd=thfd, theta_rel=np.array(rels), theta=np.array(THETA))


def main(out, part="all"):
    t0 = time.time()
    img, phi0, mask = scene()
    data = dict(np.load(out)) if os.path.exists(out) else {}
    if part in ("phi", "all"):
        data.update(phi_part(img, phi0, mask))
    if part in ("theta", "all"):
        data.update(theta_part(img, phi0, mask))
    np.savez_compressed(out, **data)
    print(f"wrote {out} in {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "config4_fd.npz"),
         sys.argv[2] if len(sys.argv) > 2 else "all")
