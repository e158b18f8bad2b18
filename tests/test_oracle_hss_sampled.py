"""CPU pinning of the oracle's randomized nested HSS construction
(oracle/hss_sampled.py, Martinsson 2011), the checker of the engine's
sampled HSS build.  Every check is independent of the construction's own
formulas: exact recovery of a matrix whose off-diagonal block rows have
known rank, the expanded bases spanning their whole off-diagonal block rows
(nestedness, by a dense SVD of each block row), the compression error
against the true matrix on the benchmark kernels, the trivial tree, and two
planted mistakes that these checks must catch."""
import numpy as np
import pytest

from oracle import hss_sampled, kernels


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def gaussian_pair(n, s, seed):
    g = np.random.default_rng(seed)
    return g.standard_normal((n, s)), g.standard_normal((n, s))


def low_rank_plus_diag(n, r, seed):
    g = np.random.default_rng(seed)
    u, v = g.standard_normal((n, r)), g.standard_normal((n, r))
    return u @ v.T + n * np.eye(n)


@pytest.mark.parametrize("n,leaf", [(1024, 64), (1000, 60)])
def test_exact_low_rank_recovered(n, leaf):
    """Off-diagonal block rows of rank 3: tolerance mode finds rank 3 at
    every node and the operator equals A to rounding."""
    a = low_rank_plus_diag(n, 3, 1)
    om, ps = gaussian_pair(n, 13, 2)
    h = hss_sampled.HSSSampled(rank=0, tol=1e-10).build_from_dense(a, leaf, om, ps)
    assert all(r == 3 for rs in h.ranks_by_depth().values() for r in rs)
    assert rel(h.to_dense(), a) < 1e-13


def test_full_bases_span_block_rows():
    """Nestedness: the expanded basis U_p = diag(U_a, U_b) R_p of every node
    spans its whole off-diagonal block row A(I_p, I_p^c) (and V_p the block
    column), checked against a dense SVD of each block row."""
    n, leaf = 1024, 64
    a = low_rank_plus_diag(n, 4, 3)
    om, ps = gaussian_pair(n, 14, 4)
    h = hss_sampled.HSSSampled(rank=0, tol=1e-10).build_from_dense(a, leaf, om, ps)
    for i, nd in enumerate(h.nodes):
        if nd.depth == 0:
            continue
        rows = np.arange(nd.start, nd.start + nd.len)
        rest = np.setdiff1d(np.arange(n), rows)
        blk_r, blk_c = a[np.ix_(rows, rest)], a[np.ix_(rest, rows)].T
        assert np.linalg.matrix_rank(blk_r) == h.U[i].shape[1]
        u, v = h.U[i], h.V[i]
        assert np.linalg.norm(u.T @ u - np.eye(u.shape[1])) < 1e-12
        assert np.linalg.norm(blk_r - u @ (u.T @ blk_r)) < 1e-12 * np.linalg.norm(blk_r)
        assert np.linalg.norm(blk_c - v @ (v.T @ blk_c)) < 1e-12 * np.linalg.norm(blk_c)


def test_cauchy_benchmark_kernel_compresses():
    """1/(x_i - y_j) on the benchmark grid, leaf 256, s = 58: tolerance 1e-10
    reproduces A to 1e-10 with ranks below the sample width; fixed rank 48
    to rounding (the block rows' numerical ranks are below 48)."""
    n = 4096
    x, y = kernels.cauchy_grid_points(n)
    a = kernels.cauchy_points(x, y)
    om, ps = gaussian_pair(n, 58, 5)
    ht = hss_sampled.HSSSampled(rank=0, tol=1e-10).build_from_dense(a, 256, om, ps)
    assert rel(ht.to_dense(), a) < 1e-10
    assert max(max(r) for r in ht.ranks_by_depth().values()) < 48
    hf = hss_sampled.HSSSampled(rank=48).build_from_dense(a, 256, om, ps)
    assert rel(hf.to_dense(), a) < 1e-13


def test_gaussian_config_kernel_compresses():
    n = 4096
    a = kernels.gaussian_points(kernels.sorted_uniform_points(n, 6), 0.05, 1e-2)
    om, ps = gaussian_pair(n, 42, 7)
    h = hss_sampled.HSSSampled(rank=32).build_from_dense(a, 128, om, ps)
    assert rel(h.to_dense(), a) < 1e-13


def test_trivial_tree_is_the_matrix():
    a = low_rank_plus_diag(200, 2, 8)
    om, ps = gaussian_pair(200, 8, 9)
    h = hss_sampled.HSSSampled(rank=4).build_from_dense(a, 256, om, ps)
    assert np.array_equal(h.to_dense(), a)


def test_planted_mistakes_are_caught():
    """Leaving out the sibling terms of the parent samples, or swapping the
    root's two couplings, must fail the dense check that the correct
    construction passes."""
    n, leaf = 1024, 64
    x, y = kernels.cauchy_grid_points(n)
    a = kernels.cauchy_points(x, y)
    om, ps = gaussian_pair(n, 40, 10)
    good = hss_sampled.HSSSampled(rank=30).build_from_dense(a, leaf, om, ps)
    assert rel(good.to_dense(), a) < 1e-10
    bad = hss_sampled.HSSSampled(rank=30).build_from_dense(a, leaf, om, ps, subtract_siblings=False)
    assert rel(bad.to_dense(), a) > 1e-6
    bab, bba = good.B[0]
    good.B[0] = (bba, bab)
    assert rel(good.to_dense(), a) > 1e-2
