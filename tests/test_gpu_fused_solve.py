"""The fused apply-inverse (one persistent launch per four columns) against
the multi-launch apply on the same factorisation: equal to rounding times
the condition number (only the summation order differs) for FP64 and mixed
precision, uneven trees, leaves of up to 512 rows (two rows per thread),
1..16 columns, the transposed solve; bit-identical from call to call.  A
spin limit too small for any wait to succeed makes the solve report
HMAT_ERR_DEVICE instead of hanging, and the next solve is unaffected."""
import numpy as np
import pytest

import hmat_amd
from oracle import kernels

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _handle(n, leaf, mixed, kind="cauchy", rank=32):
    if kind == "cauchy":
        xc, yc = kernels.cauchy_grid_points(n)
        k = hmat_amd.make_kernel(hmat_amd.KERNEL_CAUCHY, xc, yc)
    else:
        k = hmat_amd.make_kernel(hmat_amd.KERNEL_GAUSSIAN, kernels.sorted_uniform_points(n, 3), p0=0.05, p1=1e-2)
    o = hmat_amd.make_options(0.0, rank=rank, fixed_rank=True, oversample=10, power_iters=1, exact_below=0,
                              precision=hmat_amd.PREC_MIXED if mixed else hmat_amd.PREC_FP64,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=9)
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.build_kernel(k, n, leaf)
    return h


@pytest.mark.parametrize("n,leaf,mixed,kind", [(8192, 256, False, "cauchy"), (8192, 256, True, "cauchy"),
                                               (5000, 100, False, "gauss"), (3001, 64, True, "cauchy"),
                                               (6000, 512, False, "gauss")])
def test_fused_matches_multilaunch(n, leaf, mixed, kind):
    h = _handle(n, leaf, mixed, kind)
    if mixed:  # compare the apply itself: one FP32 solve, no refinement
        h.set_refinement(1.0, 0)
    # both applies are backward stable to rounding: their difference is
    # rounding times kappa (1e-7 measured in FP32 on the Cauchy system)
    bar = 1e-5 if mixed else 50 * 1.1e-16 * np.linalg.cond(h.to_dense(), 1)
    rng = np.random.default_rng(5)
    for ncols in (1, 3, 5, 16):
        b = rng.standard_normal((n, ncols))
        h.set_fused_solve(False)
        xm = h.solve(b)
        h.set_fused_solve(2)  # fused for every column count (four per launch)
        xf = h.solve(b)
        xf2 = h.solve(b)
        e = rel(xf, xm)
        print(f"\nN={n} leaf={leaf} {'mixed' if mixed else 'fp64'} {kind} {ncols} cols: fused vs multi-launch {e:.1e}",
              flush=True)
        assert np.array_equal(xf, xf2)
        assert e < bar
    # the transposed solve keeps its own factorisation and plan
    b = rng.standard_normal((n, 2))
    h.set_fused_solve(False)
    tm = h.solve_transpose(b)
    h.set_fused_solve(2)
    assert rel(h.solve_transpose(b), tm) < bar


def test_fused_fp64_solves_dense_system():
    """The fused FP64 solve against dense LAPACK on the compressed operator."""
    n = 4096
    h = _handle(n, 128, False, "gauss")
    ad = h.to_dense()
    b = np.random.default_rng(2).standard_normal((n, 3))
    x = h.solve(b)
    assert rel(x, np.linalg.solve(ad, b)) < 1e-11


def test_fused_wait_limit_reports_error():
    n = 8192
    h = _handle(n, 128, False)
    b = np.random.default_rng(1).standard_normal((n, 1))
    x0 = h.solve(b)
    hmat_amd.set_fused_spin_limit(1)
    try:
        with pytest.raises(hmat_amd.HmatError) as e:
            h.solve(b)
        assert e.value.status == 2  # HMAT_ERR_DEVICE
    finally:
        hmat_amd.set_fused_spin_limit(0)
    assert np.array_equal(h.solve(b), x0)


def test_default_mode_dispatch():
    """Mode 1 (default) takes the fused kernel for a few columns over a few
    leaves per workgroup and the multi-launch apply for more than four
    columns: bit-identical to the forced path in each case."""
    n = 8192
    h = _handle(n, 128, False)
    rng = np.random.default_rng(3)
    b1, b16 = rng.standard_normal((n, 1)), rng.standard_normal((n, 16))
    h.set_fused_solve(2)
    f1 = h.solve(b1)
    h.set_fused_solve(0)
    m16 = h.solve(b16)
    h.set_fused_solve(1)
    assert np.array_equal(h.solve(b1), f1)
    assert np.array_equal(h.solve(b16), m16)
    with pytest.raises(hmat_amd.HmatError):
        h.set_fused_solve(3)
