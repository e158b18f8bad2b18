"""Differentiable HODLR solve for Gaussian-kernel systems (torch wrapper).

    x = A(sigma, lam)^{-1} b,   A_ij = exp(-(p_i - p_j)^2 / (2 sigma^2)) + lam delta_ij

Forward: matrix-free HODLR build on the device points and a solve (FP64, or
mixed precision with FP64 refinement).  Backward, for an upstream gradient
g = dl/dx:

    w = A_H^{-T} g                       (transpose solve, its own factorisation)
    dl/db     = w
    dl/dlam   = -w^T x                    (dA/dlam = I)
    dl/dsigma = -w^T (dA/dsigma) x        (dA/dsigma_ij = exp(-d^2/(2 sigma^2)) d^2 / sigma^3,
                                           applied on the fly, FP64, never stored)

The reference has no adjoint; the solves are the engine's HODLR solves and
the derivative product is an exact dense kernel product evaluated in tiles.
The engine runs on its own HIP stream and synchronises at the end of every
call; inputs made by torch are synchronised before each call.
"""
import torch

from . import (HODLRMatrixOMP, KERNEL_GAUSSIAN, KERNEL_GAUSSIAN_DSIGMA, kernel_matvec_device,
               make_kernel_device)


def _colmajor(t):
    """(N,) or (N, k) tensor -> (storage, nrhs) with column-major storage."""
    if t.dim() == 1:
        return t.contiguous(), 1
    return t.t().contiguous(), t.shape[1]


def _from_colmajor(buf, shape):
    return buf if len(shape) == 1 else buf.t()


class GaussianHodlrSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, b, sigma, lam, points, opts, leaf):
        n = points.shape[0]
        bs, nrhs = _colmajor(b.detach().to(torch.float64))
        pts = points.detach().to(torch.float64).contiguous()
        h = HODLRMatrixOMP(opts=opts)
        k = make_kernel_device(KERNEL_GAUSSIAN, pts.data_ptr(), p0=float(sigma), p1=float(lam))
        torch.cuda.synchronize()
        h.build_kernel_device(k, n, leaf)
        xs = torch.empty_like(bs)
        h.solve_device(bs.data_ptr(), n, nrhs, xs.data_ptr(), n)
        x = _from_colmajor(xs, b.shape)
        ctx.h, ctx.n, ctx.nrhs, ctx.shape = h, n, nrhs, b.shape
        ctx.devs = (sigma.device, lam.device)
        ctx.save_for_backward(xs, pts, torch.as_tensor(sigma, dtype=torch.float64))
        return x

    @staticmethod
    def backward(ctx, gx):
        xs, pts, sigma = ctx.saved_tensors
        n, nrhs = ctx.n, ctx.nrhs
        gs, _ = _colmajor(gx.detach().to(torch.float64))
        ws = torch.empty_like(gs)
        torch.cuda.synchronize()
        ctx.h.solve_transpose_device(gs.data_ptr(), n, nrhs, ws.data_ptr(), n)
        kd = make_kernel_device(KERNEL_GAUSSIAN_DSIGMA, pts.data_ptr(), p0=float(sigma))
        ys = torch.empty_like(xs)
        kernel_matvec_device(kd, n, xs.data_ptr(), n, nrhs, ys.data_ptr(), n)
        g_b = _from_colmajor(ws, ctx.shape)
        g_sigma = (-(ws * ys).sum()).to(ctx.devs[0])
        g_lam = (-(ws * xs).sum()).to(ctx.devs[1])
        return g_b, g_sigma, g_lam, None, None, None


def gaussian_solve(b, sigma, lam, points, opts, leaf=128):
    """Differentiable x = A(sigma, lam)^{-1} b (see module docstring).

    b: (N,) or (N, k) float64 CUDA tensor; sigma, lam: 0-d tensors or
    floats; points: (N,) float64 CUDA tensor (sorted for good compression);
    opts: hmat_amd.Options (see hmat_amd.make_options)."""
    if not torch.is_tensor(sigma):
        sigma = torch.tensor(float(sigma), dtype=torch.float64)
    if not torch.is_tensor(lam):
        lam = torch.tensor(float(lam), dtype=torch.float64)
    return GaussianHodlrSolve.apply(b, sigma, lam, points, opts, leaf)
