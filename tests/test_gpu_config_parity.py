"""Same-Omega parity with the oracle on the kernels of the benchmark
configurations, at N = 4096 (sizes the oracle finishes in seconds).

  * Cauchy system A_ij = 1/(x_i - y_j), x_i = i/N, y_j = (j+1/2)/N, leaf 256,
    fixed rank 48, oversample 10, q = 2 (the precision benchmark's kernel
    and options);
  * Gaussian A_ij = exp(-(p_i-p_j)^2 / (2 sigma^2)) + lam delta_ij,
    sigma = 0.05, lam = 1e-2, sorted U[0,1] points, leaf 128, fixed rank 32,
    oversample 10, q = 2, 16 right-hand sides (the reference-comparison
    workload's kernel and options).

The GPU build takes the oracle's Omega in the serial draw order
(hodlr_routines.h:43-53, hss_routines.h:57-77).  Bars: identical per-block
ranks, the compressed operators equal to 1e-12, matvecs equal to 1e-12.  The
solutions: FP64 HODLR against the oracle's recursive Woodbury solve
(hodlr_routines_omp.h:147-240) within the rounding of the two Woodbury forms
times kappa; reference-mode HSS (identity transfers) against dense LAPACK on
the oracle's own operator (the push-through solve of hss_routines.h:197-290
is the same algebra; its column-by-column recursion over a 1024-wide root
basis would take minutes in the oracle).
"""
import numpy as np
import pytest

import hmat_amd
from oracle import hodlr as oh, hss as os_, kernels, lowrank, rng

pytestmark = pytest.mark.gpu

N = 4096


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _cauchy():
    xc, yc = kernels.cauchy_grid_points(N)
    return kernels.cauchy_points(xc, yc), 256, 48


def _gauss():
    p = kernels.sorted_uniform_points(N, 65)
    return kernels.gaussian_points(p, 0.05, 1e-2), 128, 32


def _opts(rank):
    return (lowrank.Options(tol=0.0, rank=rank, fixed_rank=True, oversample=10, power_iters=2, exact_below=0),
            hmat_amd.make_options(0.0, rank=rank, fixed_rank=True, oversample=10, power_iters=2, exact_below=0))


def _oracle_ranks(h, depth):
    return [r for pair in h.ranks_by_level()[depth - 1] for r in pair]


@pytest.mark.parametrize("case", ["cauchy_leaf256_rank48", "gaussian_s005_leaf128_rank32"])
def test_hodlr_same_omega(case):
    a, leaf, rank = _cauchy() if case.startswith("cauchy") else _gauss()
    oo, og = _opts(rank)
    s = rng.RandomStream(31, record=True)
    ho = oh.HODLRMatrix(opts=oo)
    ho.build_from_dense(a, leaf, s)
    omega = s.flat_log()
    hg = hmat_amd.HODLRMatrixOMP(opts=og)
    assert hg.omega_length(N, leaf) == omega.size
    hg.build_from_dense(a, leaf, omega=omega)
    for d in range(1, hg.levels() + 1):
        assert hg.level_ranks(d) == _oracle_ranks(ho, d), d
    G, O = hg.to_dense(), ho.to_dense()
    e_op, e_comp = rel(G, O), rel(O, a)
    nrhs = 16 if case.startswith("gaussian") else 1
    b = np.random.default_rng(3).standard_normal((N, nrhs))
    e_mv = rel(hg.multiply(b), ho.multiply_multi(b))
    xg, xo = hg.solve(b), ho.solve(b)
    xd = np.linalg.solve(O, b)  # dense LAPACK on the oracle's operator
    e_sv, e_gd, e_od = rel(xg, xo), rel(xg, xd), rel(xo, xd)
    kappa = np.linalg.cond(O, 1)
    print(f"\n{case}: operator vs oracle {e_op:.1e} (oracle vs A {e_comp:.1e}), matvec {e_mv:.1e}, "
          f"solve: GPU vs oracle {e_sv:.1e}, GPU vs dense {e_gd:.1e}, oracle vs dense {e_od:.1e}, "
          f"kappa_1 {kappa:.1e}", flush=True)
    assert e_op < 1e-12
    assert e_mv < 1e-12
    # the GPU solve is a backward-stable-order solve of the oracle's operator
    assert e_gd < 1e-14 * kappa
    # and within ten times the compression error of the oracle's own
    # recursive Woodbury solution (the drop-in bar, with the operator's
    # compression error as the tolerance of a fixed-rank build)
    assert e_sv <= 10 * max(e_comp, e_od)


@pytest.mark.parametrize("case", ["cauchy_leaf256_rank48", "gaussian_s005_leaf128_rank32"])
def test_hss_reference_mode_same_omega(case):
    a, leaf, rank = _cauchy() if case.startswith("cauchy") else _gauss()
    oo, og = _opts(rank)
    s = rng.RandomStream(41, record=True)
    ho = os_.HSSMatrix(opts=oo)
    ho.build_from_dense(a, leaf, s)
    omega = s.flat_log()
    hg = hmat_amd.HSSMatrixOMP(opts=og)
    assert hg.omega_length(N, leaf) == omega.size
    hg.build_from_dense(a, leaf, omega=omega)
    assert hg.root_rank() == ho.root.rank_U
    G, O = hg.to_dense(), ho.to_dense()
    e_op = rel(G, O)
    nrhs = 16 if case.startswith("gaussian") else 1
    b = np.random.default_rng(4).standard_normal((N, nrhs))
    e_mv = rel(hg.multiply(b), ho.multiply_multi(b))
    kappa = np.linalg.cond(O, 1)
    xo = np.linalg.solve(O, b)
    # the HSS solve takes every column through the C ABI (the reference OMP
    # class solves column 0; the serial class loops over columns)
    xg = hg.solve(b)
    e_sv = rel(xg, xo)
    print(f"\nHSS {case}: root rank {hg.root_rank()}, operator vs oracle {e_op:.1e}, matvec {e_mv:.1e}, "
          f"solve vs dense LAPACK of the oracle's operator {e_sv:.1e}, kappa {kappa:.1e}", flush=True)
    assert e_op < 1e-12
    assert e_mv < 1e-12
    # push-through Woodbury amplifies rounding through its capacitance
    # systems; refinement (below) removes it
    assert e_sv < 1e-13 * kappa * 100
    hg.set_refinement(1e-12, 10)
    xr = hg.solve(b)
    e_r = rel(xr, xo)
    print(f"HSS {case}: refined: resid {hg.last_refinement()}, vs dense {e_r:.1e}", flush=True)
    assert rel(G @ xr, b) <= 1.5e-12
    assert e_r < 1e-15 * kappa * 100
