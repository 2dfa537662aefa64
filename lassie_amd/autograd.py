"""torch.autograd front end of the differentiation form.

``trajectories(log_k, models, t, **opts)`` integrates B parameterisations of
one reaction network (B uploaded models with the same species and reactions)
with rate constants exp(log_k[b]) and returns their sampled trajectories
[B, nt, S]; backpropagating through it runs the discrete adjoint
(lassie_adjoint_batch) through the recorded RKF45/BDF1 steps.  A
parameterisation whose trajectory contains a BDF1 step whose modified Newton
iteration did not converge has no finite gradient; backward then raises
LassieError (LASSIE_ERR_NONFINITE) rather than hand NaNs to an optimizer.

Tensors cross the boundary as FP64 host arrays: device memory belongs to the
library (the forward pass, its step record and the reverse pass all stay on
the GPU), so torch is used here only for autograd bookkeeping.
"""
import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def _function():
    torch = _torch()

    class LassieTrajectories(torch.autograd.Function):
        @staticmethod
        def forward(ctx, log_k, models, t, max_steps, opts):
            lk = log_k.detach().cpu().numpy().astype(np.float64)
            if lk.shape[0] != len(models):
                raise ValueError("one row of log_k per model")
            for b, dm in enumerate(models):
                dm.set_rates(np.exp(lk[b]))
            outs, stats, ms, tape = _lib.forward_batch(models, t, max_steps=max_steps, **opts)
            ctx.tape = tape
            ctx.stats = stats
            return torch.from_numpy(np.stack(outs))

        @staticmethod
        def backward(ctx, grad_out):
            g = grad_out.detach().cpu().numpy().astype(np.float64)
            try:
                gks, _, _ = ctx.tape.adjoint(list(g))
            finally:
                ctx.tape.close()
            return torch.from_numpy(np.stack(gks)), None, None, None, None

    return LassieTrajectories


_FN = None


def trajectories(log_k, models, t, max_steps=1 << 16, **opts):
    """Sampled states [B, nt, S] of B models with k = exp(log_k[b]);
    differentiable with respect to log_k (float64 tensor [B, R])."""
    global _FN
    if _FN is None:
        _FN = _function()
    if any(dm.S != models[0].S or dm.R != models[0].R for dm in models):
        raise ValueError("the parameterisations must share one network")
    return _FN.apply(log_k, list(models), np.ascontiguousarray(t, np.float64), int(max_steps), dict(opts))


def log_loss(traj, target, eps=1e-12):
    """sum over samples and species of (log(y + eps) - log(y* + eps))^2."""
    torch = _torch()
    return torch.sum((torch.log(traj + eps) - torch.log(target + eps)) ** 2)
