"""torch.autograd wrappers of the GPU factorization (thin: the work is in the C ABI).

``logdet_quad(values, b, factor)`` returns logdet(A) + b^T A^{-1} b for the SPD
matrix whose stored lower-triangle values (input order of the analysed pattern)
are ``values``; its backward gives d/dvalues through the supernodal selected
inverse (swc_gradient) and d/db = 2 A^{-1} b.  Each stored off-diagonal value
a_ij stands for both a_ij and a_ji, so its gradient counts both.
"""
from __future__ import annotations

import torch

from .api import Factor


def _sync():
    if torch.cuda.is_available():
        torch.cuda.current_stream().synchronize()


class _LogdetQuad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, b, factor: Factor, alpha: float, beta: float):
        vals = values.detach().to(torch.float64).contiguous()
        rhs = b.detach().to(torch.float64).contiguous()
        dev = vals.device
        _sync()
        info = factor.factor(vals if vals.is_cuda else vals.numpy())
        if info:
            raise RuntimeError(f"matrix is not positive definite (column {info - 1})")
        x = torch.from_numpy(factor.solve(rhs.cpu().numpy())).to(dev)
        out = alpha * factor.logdet() + beta * float(torch.dot(rhs.cpu(), x.cpu()))
        ctx.factor, ctx.alpha, ctx.beta = factor, alpha, beta
        ctx.save_for_backward(x)
        return torch.tensor(out, dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        f = ctx.factor
        _sync()
        if x.is_cuda:
            gv = torch.empty(f.bcsc.nnz, dtype=torch.float64, device=x.device)
            f.gradient(x, alpha=ctx.alpha, beta=ctx.beta, out=gv)
        else:
            gv = torch.from_numpy(f.gradient(x.numpy(), alpha=ctx.alpha, beta=ctx.beta))
        return g * gv, g * 2.0 * ctx.beta * x, None, None, None


def logdet_quad(values, b, factor: Factor, alpha=1.0, beta=1.0):
    """alpha logdet(A) + beta b^T A^{-1} b, differentiable in values and b."""
    return _LogdetQuad.apply(values, b, factor, float(alpha), float(beta))
