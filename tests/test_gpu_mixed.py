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
    assert e_mixed < 1e-5


def test_mixed_operator_close_to_fp64_operator():
    n = 4096
    hm, a = _cauchy(n, hmat_amd.PREC_MIXED)
    hf, _ = _cauchy(n, hmat_amd.PREC_FP64)
    Am, Af = hm.to_dense(), hf.to_dense()
    print(f"\noperator: mixed vs A {rel(Am, a):.2e}, fp64 vs A {rel(Af, a):.2e}")
    assert rel(Af, a) < 1e-12
    assert rel(Am, a) < 1e-6


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
    assert rel(Am, a) < 1e-6
    b = np.random.default_rng(2).standard_normal((n, 1))
    x = hm.solve(b)
    it, res = hm.last_refinement()
    assert res <= 1e-10 and 1 <= it <= 10
    assert rel(Am @ x, b) <= 1.5e-10
