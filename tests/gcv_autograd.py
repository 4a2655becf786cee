"""torch.autograd view of the library's hand-written GCV adjoint (test harness).

`gcv_score(log_sp, data, targets, S)` returns the mgcv GCV scores of a batch of
regrXonS fits (all sharing |S|) at log smoothing parameters `log_sp` (count x
m, float64).  Its backward is the gradient from kpc_gcv_grad_hess, and the
backward of that gradient is the Hessian from the same call, so
torch.autograd.grad(..., create_graph=True) and
torch.autograd.functional.hessian see exactly the library's derivatives --
no finite differences and no torch-side re-derivation.  Only the GPU library
is evaluated; the functions here are wrappers.
"""
import numpy as np
import torch

import fastkpc_amd as fk


def _eval(log_sp, data, targets, S):
    x = np.ascontiguousarray(log_sp.detach().cpu().numpy(), dtype=np.float64)
    return fk.gcv_grad_hess(data, targets, S, x)


class _GCVGradient(torch.autograd.Function):
    """(count, m) gradient; its backward applies the (symmetric) Hessian."""

    @staticmethod
    def forward(ctx, log_sp, data, targets, S):
        _, g, H = _eval(log_sp, data, targets, S)
        ctx.save_for_backward(torch.from_numpy(H))
        return torch.from_numpy(g).to(log_sp)

    @staticmethod
    def backward(ctx, gout):
        (H,) = ctx.saved_tensors
        return (H.to(gout) @ gout[:, :, None])[:, :, 0], None, None, None


class GCVScore(torch.autograd.Function):
    """(count,) GCV scores at log_sp (count, m)."""

    @staticmethod
    def forward(ctx, log_sp, data, targets, S):
        score, _, _ = _eval(log_sp, data, targets, S)
        ctx.save_for_backward(log_sp)
        ctx.args = (data, targets, S)
        return torch.from_numpy(score).to(log_sp)

    @staticmethod
    def backward(ctx, gout):
        (log_sp,) = ctx.saved_tensors
        g = _GCVGradient.apply(log_sp, *ctx.args)
        return gout[:, None] * g, None, None, None


def gcv_score(log_sp, data, targets, S):
    return GCVScore.apply(log_sp, data, targets, S)


def newton(log_sp0, data, targets, S, iterations=40, max_step=5.0, halvings=25, tol=1e-10):
    """Batched damped Newton on the GCV score in log(sp), every derivative
    taken through autograd (the library's adjoint): per fit, the Newton step
    (steepest descent where the Hessian is not positive definite), capped at
    max_step, halved until the score decreases."""
    x = log_sp0.detach().clone()
    count, m = x.shape
    for _ in range(iterations):
        xv = x.clone().requires_grad_(True)
        f = gcv_score(xv, data, targets, S)
        (g,) = torch.autograd.grad(f.sum(), xv, create_graph=True)
        H = torch.stack([torch.autograd.grad(g[:, j].sum(), xv, retain_graph=True)[0]
                         for j in range(m)], dim=1)
        g = g.detach()
        step = torch.empty_like(x)
        for c in range(count):
            w, V = torch.linalg.eigh(H[c])
            if bool((w > 0).all()):
                step[c] = -(V @ ((V.T @ g[c]) / w))
            else:
                step[c] = -g[c] / g[c].abs().max().clamp_min(1e-300)
        big = step.abs().max(dim=1).values
        step = torch.where((big > max_step)[:, None], step * (max_step / big)[:, None], step)
        f0 = f.detach()
        done = torch.zeros(count, dtype=torch.bool)
        for _ in range(halvings):
            trial = x + step
            ft = gcv_score(trial, data, targets, S).detach()
            better = (ft < f0) & ~done
            x = torch.where(better[:, None], trial, x)
            done |= better
            if bool(done.all()):
                break
            step = torch.where(done[:, None], step, step / 2)
        if bool((g.abs().max(dim=1).values <= tol * (1 + f0.abs())).all()):
            break
    return x
