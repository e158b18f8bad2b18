"""Mixed precision (BF16-split range finder, FP32 factorisation, FP64
projection and refinement) against the plain FP64 path and dense LAPACK."""
import numpy as np
import pytest

import hmat_amd
from oracle import kernels

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _cauchy(n, prec, q=2, seed=0x5EED):
    xc, yc = kernels.cauchy_grid_points(n)
    o = hmat_amd.make_options(0.0, rank=48, fixed_rank=True, oversample=10, power_iters=q,
                              exact_below=0, precision=prec, omega_dist=hmat_amd.OMEGA_GAUSSIAN,
                              seed=seed)
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.build_kernel(hmat_amd.make_kernel(hmat_amd.KERNEL_CAUCHY, xc, yc), n, 256)
    return h, kernels.cauchy_points(xc, yc)


def test_mixed_refines_cauchy_to_1e10():
    n = 8192
    hm, a = _cauchy(n, hmat_amd.PREC_MIXED)
    hf, _ = _cauchy(n, hmat_amd.PREC_FP64)
    b = np.random.default_rng(0).standard_normal((n, 2))
    xm = hm.solve(b)
    it, res = hm.last_refinement()
    assert res <= 1e-10 and 1 <= it <= 10
    # residual against the mixed operator, recomputed on the host side
    assert rel(hm.multiply(xm), b) <= 1.5e-10
    xf = hf.solve(b)
    xd = np.linalg.solve(a, b)
    e_mixed, e_fp64 = rel(xm, xd), rel(xf, xd)
    print(f"\nN={n}: mixed vs dense {e_mixed:.2e}, fp64 vs dense {e_fp64:.2e}, "
          f"mixed vs fp64 {rel(xm, xf):.2e}, refinement iters {it}")
    assert e_fp64 < 1e-10
    # the BF16x3 range finder resolves blocks to ~1e-8 of their norm
    assert e_mixed < 1e-7
    assert rel(xm, xf) < 1e-7


def test_mixed_operator_close_to_fp64_operator():
    n = 4096
    hm, a = _cauchy(n, hmat_amd.PREC_MIXED)
    hf, _ = _cauchy(n, hmat_amd.PREC_FP64)
    Am, Af = hm.to_dense(), hf.to_dense()
    print(f"\noperator: mixed vs A {rel(Am, a):.2e}, fp64 vs A {rel(Af, a):.2e}")
    assert rel(Af, a) < 1e-12
    assert rel(Am, a) < 1e-7
    assert rel(Am, Af) < 1e-7


def test_mixed_reference_tolerance_mode():
    """Mixed precision with the reference options (tolerance truncation) on
    the reference driver's Cauchy matrix (kappa ~ 1.4e3): refinement reaches
    1e-10 on the compressed operator; against the true A the residual is
    bounded by the ~1e-8 accuracy of the BF16x3 range finder."""
    n = 2048
    a = kernels.generate_matrix(kernels.CAUCHY, n)
    o = hmat_amd.reference_options(1e-8)
    o.precision = hmat_amd.PREC_MIXED
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.build_from_dense(a, 256)
    b = a @ np.ones((n, 1))
    x = h.solve(b)
    it, res = h.last_refinement()
    true_res = rel(a @ x, b)
    print(f"\nref Cauchy N={n}: resid(A_H) {res:.1e} resid(A) {true_res:.1e} "
          f"err vs x_true {rel(x, np.ones((n, 1))):.1e} iters {it}")
    assert res <= 1e-10
    assert true_res < 1e-7
    assert rel(x, np.ones((n, 1))) < 1e-3


def test_mixed_ragged_tree():
    """Odd N and leaf sizes: rows of the BF16 planes start unaligned and
    reduction ranges end inside a K-tile, so the low-precision sketches stage
    their tiles element-wise, several per workgroup.  The operator stays
    within the mixed path's accuracy of the FP64 one and the refined solve
    reaches 1e-10."""
    n = 3001
    xc, yc = kernels.cauchy_grid_points(n)
    a = kernels.cauchy_points(xc, yc)

    def build(prec):
        o = hmat_amd.make_options(0.0, rank=48, fixed_rank=True, oversample=10, power_iters=2,
                                  exact_below=0, precision=prec, omega_dist=hmat_amd.OMEGA_GAUSSIAN,
                                  seed=0x5EED)
        h = hmat_amd.HODLRMatrixOMP(opts=o)
        h.build_kernel(hmat_amd.make_kernel(hmat_amd.KERNEL_CAUCHY, xc, yc), n, 100)
        return h

    hm, hf = build(hmat_amd.PREC_MIXED), build(hmat_amd.PREC_FP64)
    Am, Af = hm.to_dense(), hf.to_dense()
    print(f"\nN={n}, leaf 100: mixed vs A {rel(Am, a):.2e}, fp64 vs A {rel(Af, a):.2e}")
    assert rel(Af, a) < 1e-12
    assert rel(Am, a) < 3e-7  # measured 1.1e-7
    b = np.random.default_rng(2).standard_normal((n, 1))
    x = hm.solve(b)
    it, res = hm.last_refinement()
    assert res <= 1e-10 and 1 <= it <= 10
    assert rel(Am @ x, b) <= 1.5e-10


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_cfg3_full_size(mode):
    """The precision benchmark's own size: N = 2^18 Cauchy system, leaf 256,
    fixed rank 48, p 10, q 2, one right-hand side b ~ N(0,1), refinement to
    1e-10 through the FP64 compressed operator.  The true-matrix residual
    ||b - A x|| / ||b|| uses the exact FP64 product (every entry generated on
    the device, A never stored; the mixed path is bounded by its range
    finder's ~1e-8 accuracy), and the mixed solution is compared with the
    FP64 one."""
    import torch
    n = 1 << 18
    xs = torch.arange(n, dtype=torch.float64, device="cuda") / n
    ys = (torch.arange(n, dtype=torch.float64, device="cuda") + 0.5) / n
    b = torch.tensor(np.random.default_rng(0).standard_normal(n), device="cuda")
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_CAUCHY, xs.data_ptr(), ys.data_ptr())

    def run(prec):
        o = hmat_amd.make_options(0.0, rank=48, fixed_rank=True, oversample=10, power_iters=2,
                                  exact_below=0, precision=prec, omega_dist=hmat_amd.OMEGA_GAUSSIAN,
                                  seed=0x5EED)
        h = hmat_amd.HODLRMatrixOMP(opts=o)
        h.set_refinement(1e-10, 30)
        torch.cuda.synchronize()
        h.build_kernel_device(k, n, 256)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        h.solve_device(b.data_ptr(), n, 1, x.data_ptr(), n)
        torch.cuda.synchronize()
        return h, x

    def true_resid(x):
        y = torch.zeros_like(x)
        hmat_amd.kernel_matvec_device(k, n, x.data_ptr(), n, 1, y.data_ptr(), n)
        torch.cuda.synchronize()
        return float(torch.linalg.norm(b - y) / torch.linalg.norm(b))

    hf, xf = run(hmat_amd.PREC_FP64)
    rf = true_resid(xf)
    itf, resf = hf.last_refinement()
    if mode == "fp64":
        print(f"\nN=2^18 FP64: resid(A_H) {resf:.1e} ({itf} steps), resid(A) {rf:.1e}", flush=True)
        assert hf.levels() == 10
        assert resf <= 1e-10
        assert rf < 1e-13
        return
    del hf
    hm, xm = run(hmat_amd.PREC_MIXED)
    itm, resm = hm.last_refinement()
    rm = true_resid(xm)
    e = float(torch.linalg.norm(xm - xf) / torch.linalg.norm(xf))
    print(f"\nN=2^18 mixed: resid(A_H) {resm:.1e} ({itm} steps), resid(A) {rm:.1e}, vs FP64 {e:.1e}",
          flush=True)
    assert resm <= 1e-10 and 1 <= itm <= 5
    assert rm < 1e-7
    assert e < 1e-7
