"""Differentiable RSF: K unrolled steps as a torch.autograd.Function.

The forward runs K explicit Euler steps of the device engine (one batched
launch per sweep for all blocks) and keeps phi at every `checkpoint_every`-th
step.  The backward walks the segments in reverse: it recomputes the states of
one segment from its checkpoint and applies the hand-written adjoint of one
step (rsf_step_adjoint_device) from the last state of the segment to the first,
accumulating dLoss/d(lambda1, lambda2, nu, mu) on the device.

Gradients flow to phi0 (e.g. the output of a U-Net seed producer,
Unet/run_model.py) and to theta = (lambda1, lambda2, nu, mu); the image, sigma,
dt and epsilon are constants.
"""
from __future__ import annotations

import torch

from . import MODE_FAST, Context, make_params


class _Unrolled(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi0, theta, image, blocks, rsf, K, checkpoint_every, fixed):
        th = [float(v) for v in theta.detach().cpu().tolist()]
        p = make_params(lambda1=th[0], lambda2=th[1], nu=th[2], mu=th[3], iterations=1, **fixed)
        cur = phi0.detach().contiguous().clone()
        ckpts = [cur.clone()]
        rsf.hold_image(image)   # the image is constant over the unrolled steps
        try:
            # one graph-replayed run per checkpoint segment, in place: the same kernels
            # as step by step (the backward recomputes the segments with step_blocks),
            # without a transform and a host round trip per step
            for s in range(0, K, checkpoint_every):
                n = min(checkpoint_every, K - s)
                seg = make_params(lambda1=th[0], lambda2=th[1], nu=th[2], mu=th[3], iterations=n, **fixed)
                rsf.run_batched(blocks, image, cur, seg)
                if s + n < K:
                    ckpts.append(cur.clone())
        finally:
            rsf.release_image()
        ctx.rsf, ctx.blocks, ctx.image, ctx.p = rsf, blocks, image, p
        ctx.K, ctx.C, ctx.ckpts = K, checkpoint_every, ckpts
        ctx.theta_device = theta.device
        return cur.clone()

    @staticmethod
    def backward(ctx, gout):
        rsf, blocks, image, p = ctx.rsf, ctx.blocks, ctx.image, ctx.p
        g = gout.detach().contiguous().clone()
        gp = torch.zeros(4, dtype=torch.float32, device=g.device)
        K, C = ctx.K, ctx.C
        starts = list(range(0, K, C))
        rsf.hold_image(image)
        try:
            for si in reversed(range(len(starts))):
                s = starts[si]
                e = min(s + C, K)
                states = [ctx.ckpts[si]]
                for _ in range(s, e - 1):
                    nxt = torch.empty_like(states[-1])
                    rsf.step_blocks(blocks, image, states[-1], nxt, p)
                    states.append(nxt)
                for t in reversed(range(s, e)):
                    gin = torch.empty_like(g)
                    rsf.adjoint(blocks, image, states[t - s], g, gin, gp, p)
                    g = gin
                del states
        finally:
            rsf.release_image()
        ctx.ckpts = None
        return g, gp.to(ctx.theta_device), None, None, None, None, None, None


def rsf_unrolled(phi0: torch.Tensor, theta: torch.Tensor, image: torch.Tensor, blocks, K: int,
                 checkpoint_every: int = 8, sigma: float = 3.0, dt: float = 0.1, epsilon: float = 1.0,
                 mode: int = MODE_FAST, context: Context | None = None) -> torch.Tensor:
    """phi_K of K RSF steps from phi0 over packed blocks [(offset, w, h, l), ...].

    phi0: flat float32 device tensor holding the packed blocks; theta: 4-element
    tensor (lambda1, lambda2, nu, mu); image: flat uint8/float32 device tensor.
    Differentiable in phi0 and theta.
    """
    if K < 1 or checkpoint_every < 1:
        raise ValueError("K and checkpoint_every must be >= 1")
    rsf = context if context is not None else _default_context(phi0.device)
    fixed = dict(sigma=sigma, dt=dt, epsilon=epsilon, mode=mode)
    return _Unrolled.apply(phi0, theta, image, blocks, rsf, int(K), int(checkpoint_every), fixed)


_contexts: dict = {}


def _default_context(device) -> Context:
    idx = device.index if device.index is not None else 0
    if idx not in _contexts:
        _contexts[idx] = Context(idx)
    return _contexts[idx]


def bce_inside_loss(phi: torch.Tensor, mask: torch.Tensor, epsilon: float = 1.0) -> torch.Tensor:
    """BCE of sigmoid(-phi/epsilon) (inside probability; phi < 0 inside) against a 0/1 mask."""
    return torch.nn.functional.binary_cross_entropy_with_logits(-phi / epsilon, mask.to(phi.dtype))
