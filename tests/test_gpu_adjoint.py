"""Transpose solve, dA/dsigma product and the differentiable Gaussian solve.

Loss l = 1/2 ||x - t||^2 with A(sigma, lam) x = b, A_ij = exp(-(p_i-p_j)^2 /
(2 sigma^2)) + lam delta_ij, sorted U[0,1] points, sigma = 0.1, lam = 1e-2,
b, t ~ N(0,1).  Bars:
  * transpose matvec / solve equal the dense transpose of the compressed
    operator to 1e-13 / 1e-10 (exact algebra);
  * the dA/dsigma product equals the closed form to 1e-13;
  * gradients (db, dsigma, dlam) vs the FP64 dense gradient at N = 4096 to
    1e-6 relative (compression tol 1e-12 times the conditioning of A);
  * dsigma and dlam vs fourth-order central differences of the engine's own
    loss (fixed rank, so the compressed operator is smooth in sigma) to 3e-5
    (the difference quotient carries the solve's rounding divided by h).
"""
import numpy as np
import pytest
import torch

import hmat_amd
from hmat_amd import autograd as had
from oracle import kernels

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_transpose_ops_exact():
    n = 3000
    xc, yc = kernels.cauchy_grid_points(n)
    a = kernels.cauchy_points(xc, yc)  # non-symmetric, well conditioned
    o = hmat_amd.make_options(1e-10, rank=32, oversample=10, power_iters=1, exact_below=0,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=3)
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.build_from_dense(a, 100)
    ad = h.to_dense()
    v = np.random.default_rng(0).standard_normal((n, 3))
    assert rel(h.multiply_transpose(v), ad.T @ v) < 1e-13
    assert rel(h.solve_transpose(v), np.linalg.solve(ad.T, v)) < 1e-10
    # the forward operator is untouched by the transposed calls
    assert rel(h.multiply(v), ad @ v) < 1e-13
    assert rel(h.solve(v), np.linalg.solve(ad, v)) < 1e-10
    assert rel(h.to_dense(), ad) == 0.0


def test_transpose_solve_mixed_precision_refines():
    n = 4096
    xc, yc = kernels.cauchy_grid_points(n)
    o = hmat_amd.make_options(0.0, rank=32, fixed_rank=True, oversample=10, power_iters=1,
                              exact_below=0, precision=hmat_amd.PREC_MIXED,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN)
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.build_kernel(hmat_amd.make_kernel(hmat_amd.KERNEL_CAUCHY, xc, yc), n, 128)
    b = np.random.default_rng(1).standard_normal((n, 1))
    w = h.solve_transpose(b)
    assert rel(h.multiply_transpose(w), b) <= 1e-10


def test_dsigma_product_closed_form():
    n = 2048
    p = np.sort(np.random.default_rng(5).random(n))
    sigma = 0.1
    d2 = (p[:, None] - p[None, :]) ** 2
    da = np.exp(-d2 / (2 * sigma * sigma)) * d2 / sigma ** 3
    x = np.random.default_rng(6).standard_normal((n, 2))
    dp = torch.tensor(p, device="cuda")
    dx = torch.tensor(np.asfortranarray(x).T.copy(), device="cuda")  # column-major storage
    dy = torch.empty_like(dx)
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_GAUSSIAN_DSIGMA, dp.data_ptr(), p0=sigma)
    torch.cuda.synchronize()
    hmat_amd.kernel_matvec_device(k, n, dx.data_ptr(), n, 2, dy.data_ptr(), n)
    assert rel(dy.cpu().numpy().T, da @ x) < 1e-13


def _dense_grads(p, sigma, lam, b, t):
    d2 = (p[:, None] - p[None, :]) ** 2
    a0 = np.exp(-d2 / (2 * sigma * sigma))
    a = a0 + lam * np.eye(len(p))
    x = np.linalg.solve(a, b)
    w = np.linalg.solve(a.T, x - t)
    return x, w, -w @ ((a0 * d2 / sigma ** 3) @ x), -w @ x


def _opts(tol=1e-12, fixed=False, rank=48):
    return hmat_amd.make_options(tol, rank=rank, fixed_rank=fixed, oversample=10, power_iters=2,
                                 exact_below=0, omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=17)


def test_gradients_against_dense_fp64():
    n, sigma, lam = 4096, 0.1, 1e-2
    rng = np.random.default_rng(2024)
    p = np.sort(rng.random(n))
    b, t = rng.standard_normal(n), rng.standard_normal(n)
    xd, wd, gsd, gld = _dense_grads(p, sigma, lam, b, t)
    dp = torch.tensor(p, device="cuda")
    db = torch.tensor(b, device="cuda", requires_grad=True)
    ts = torch.tensor(sigma, dtype=torch.float64, requires_grad=True)
    tl = torch.tensor(lam, dtype=torch.float64, requires_grad=True)
    x = had.gaussian_solve(db, ts, tl, dp, _opts(), leaf=128)
    loss = 0.5 * ((x - torch.tensor(t, device="cuda")) ** 2).sum()
    loss.backward()
    assert rel(x.detach().cpu().numpy(), xd) < 1e-6
    assert rel(db.grad.cpu().numpy(), wd) < 1e-6
    assert abs(ts.grad.item() - gsd) <= 1e-6 * abs(gsd)
    assert abs(tl.grad.item() - gld) <= 1e-6 * abs(gld)


def test_gradients_against_central_differences():
    n, sigma, lam = 8192, 0.1, 1e-2
    rng = np.random.default_rng(7)
    p = np.sort(rng.random(n))
    dp = torch.tensor(p, device="cuda")
    b = torch.tensor(rng.standard_normal(n), device="cuda")
    t = torch.tensor(rng.standard_normal(n), device="cuda")
    o = _opts(tol=0.0, fixed=True, rank=40)

    def loss(s, l):
        x = had.gaussian_solve(b, s, l, dp, o, leaf=128)
        return 0.5 * float(((x - t) ** 2).sum())

    ts = torch.tensor(sigma, dtype=torch.float64, requires_grad=True)
    tl = torch.tensor(lam, dtype=torch.float64, requires_grad=True)
    x = had.gaussian_solve(b, ts, tl, dp, o, leaf=128)
    (0.5 * ((x - t) ** 2).sum()).backward()
    # fourth-order central stencil; h balances the O(h^4) truncation against
    # the solve's rounding (kappa(A) ~ 1e5, so ~1e-11 of the loss) divided
    # by h, which limits the difference quotient to a few 1e-6 relative
    def fd(f, v, h):
        return (-f(v + 2 * h) + 8 * f(v + h) - 8 * f(v - h) + f(v - 2 * h)) / (12 * h)

    fd_s = fd(lambda s: loss(s, lam), sigma, 2e-3 * sigma)
    fd_l = fd(lambda l: loss(sigma, l), lam, 2e-3 * lam)
    assert abs(ts.grad.item() - fd_s) <= 3e-5 * abs(fd_s)
    assert abs(tl.grad.item() - fd_l) <= 3e-5 * abs(fd_l)


def test_cfg5_size_forward_backward():
    """N = 65536 (the differentiation workload): gradients finite; dlam
    against a second-order central difference and dsigma against a
    fourth-order one (the compressed operator is smooth in sigma at fixed
    rank and a fixed Omega seed)."""
    n, sigma, lam = 65536, 0.1, 1e-2
    rng = np.random.default_rng(11)
    dp = torch.tensor(np.sort(rng.random(n)), device="cuda")
    b = torch.tensor(rng.standard_normal(n), device="cuda", requires_grad=True)
    t = torch.tensor(rng.standard_normal(n), device="cuda")
    o = _opts(tol=0.0, fixed=True, rank=32)
    ts = torch.tensor(sigma, dtype=torch.float64, requires_grad=True)
    tl = torch.tensor(lam, dtype=torch.float64, requires_grad=True)
    x = had.gaussian_solve(b, ts, tl, dp, o, leaf=128)
    (0.5 * ((x - t) ** 2).sum()).backward()
    assert torch.isfinite(b.grad).all() and np.isfinite(ts.grad.item()) and np.isfinite(tl.grad.item())

    def loss(s, l):
        x = had.gaussian_solve(b.detach(), s, l, dp, o, leaf=128)
        return 0.5 * float(((x - t) ** 2).sum())

    hl = 1e-5 * lam
    fd_l = (loss(sigma, lam + hl) - loss(sigma, lam - hl)) / (2 * hl)
    hs = 2e-3 * sigma
    fd_s = (-loss(sigma + 2 * hs, lam) + 8 * loss(sigma + hs, lam) - 8 * loss(sigma - hs, lam)
            + loss(sigma - 2 * hs, lam)) / (12 * hs)
    e_l = abs(tl.grad.item() - fd_l) / abs(fd_l)
    e_s = abs(ts.grad.item() - fd_s) / abs(fd_s)
    print(f"\nN={n}: dl/dlam {tl.grad.item():.6e} vs FD {fd_l:.6e} ({e_l:.1e}); "
          f"dl/dsigma {ts.grad.item():.6e} vs FD {fd_s:.6e} ({e_s:.1e})", flush=True)
    assert e_l <= 1e-5
    assert e_s <= 1e-4
