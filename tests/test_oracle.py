"""Pin the CPU oracle (CPU only).

The reference ships no tests or golden vectors (its drivers only print
errors), so each oracle function is checked against something it does not
restate: the C++ standard's known answer, an independent MT19937, libstdc++
golden vectors, numpy's SVD/solve, closed forms, and planted mistakes that
the checks must catch.
"""
import json
import os

import numpy as np
import pytest

from oracle import hodlr, hss, kernels, lowrank, rng

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- RNG ----
def test_mt19937_known_answer():
    # C++ [rand.predef]: 10000th invocation of a default-constructed mt19937
    assert int(rng.MT19937(5489).integers(10000)[-1]) == 4123659995


def test_mt19937_matches_numpy_legacy_seeding():
    bg = np.random.MT19937()
    bg._legacy_seeding(2024)
    assert np.array_equal(rng.MT19937(2024).integers(3000), bg.random_raw(3000))


def test_uniform_stream_matches_libstdcxx_golden():
    gold = json.load(open(os.path.join(GOLDEN, "random_stream.json")))
    for seed, vals in gold.items():
        got = rng.canonical(rng.MT19937(int(seed)), len(vals))
        assert np.array_equal(got, np.array(vals)), seed


def test_matrix_random_fills_column_major():
    s = rng.RandomStream(7)
    m = s.random(3, 2)
    flat = rng.canonical(rng.MT19937(7), 6)
    assert m.flags.f_contiguous
    assert np.array_equal(m[:, 0], flat[:3]) and np.array_equal(m[:, 1], flat[3:])


# ------------------------------------------------------------ kernels ----
def test_reference_kernels_closed_form():
    n = 5
    x = np.arange(n) / (n - 1.0)
    a = kernels.generate_cauchy(n, 1.0)
    assert a[1, 3] == 1.0 / (x[1] + x[3] + 1.0)
    assert a[2, 2] == 1.0 / (2 * x[2] + 1.0) + 1.0
    g = kernels.generate_gaussian(n, 0.5, 1e-6)
    assert g[0, 4] == np.exp(-1.0 / 0.25)
    assert g[3, 3] == 1.0 + 1e-6
    c = kernels.generate_coulomb(n, 2.0)
    assert c[0, 1] == 4.0 and c[1, 1] == 2.0
    assert np.allclose(kernels.generate_matrix(kernels.COULOMB, n).diagonal(), 10.0 * n)


def test_point_kernels_closed_form():
    x = np.array([0.1, 0.3, 0.7])
    g = kernels.gaussian_points(x, 0.1, 0.5)
    assert np.isclose(g[0, 1], np.exp(-0.04 / 0.02))
    assert np.isclose(g[2, 2], 1.5)
    xc, yc = kernels.cauchy_grid_points(8)
    a = kernels.cauchy_points(xc, yc)
    assert a[3, 1] == 1.0 / ((3 - 1.5) / 8)


# -------------------------------------------------------- compression ----
def _planted(m, n, svals, seed=0):
    rs = np.random.default_rng(seed)
    u, _ = np.linalg.qr(rs.standard_normal((m, len(svals))))
    v, _ = np.linalg.qr(rs.standard_normal((n, len(svals))))
    return np.asfortranarray(u @ np.diag(svals) @ v.T)


def test_compress_svd_truncation_rule():
    s = np.array([1.0, 1e-3, 1e-7, 1e-9, 1e-12])
    a = _planted(40, 30, s)
    assert lowrank.compress_svd(a, 1e-8).rank == 3
    assert lowrank.compress_svd(a, 1e-10).rank == 4
    assert lowrank.compress_svd(a, -1.0).rank == 30        # tol=-1 keeps all
    z = lowrank.compress_svd(np.zeros((6, 4)), 1e-8)        # S0 <= 1e-15 -> rank 1
    assert z.rank == 1 and z.B[0, 0] == 0.0
    lr = lowrank.compress_svd(a, 1e-8)
    np.testing.assert_allclose(lr.U.T @ lr.U, np.eye(3), atol=1e-13)
    np.testing.assert_allclose(lr.V.T @ lr.V, np.eye(3), atol=1e-13)
    # Eckart-Young: spectral error of the rank-3 truncation is S_3
    err = np.linalg.norm(a - lr.U @ lr.B @ lr.V.T, 2)
    assert abs(err - 1e-9) < 1e-12


def test_compress_svd_matches_numpy_gesdd():
    a = np.asfortranarray(np.random.default_rng(3).standard_normal((50, 20)))
    lr = lowrank.compress_svd(a, -1.0)
    np.testing.assert_allclose(np.diag(lr.B), np.linalg.svd(a, compute_uv=False), rtol=1e-13)


def test_compress_rsvd_fallback_below_256():
    a = _planted(200, 300, np.logspace(0, -12, 10))
    s1 = rng.RandomStream(1, record=True)
    lr = lowrank.compress_rsvd(a, 1e-8, 64, s1)
    assert s1.log == []  # no Omega drawn
    assert lr.rank == lowrank.compress_svd(a, 1e-8).rank


def test_compress_rsvd_accuracy_and_draw_shape():
    svals = np.logspace(0, -14, 30)
    a = _planted(400, 300, svals, seed=5)
    s = rng.RandomStream(11, record=True)
    lr = lowrank.compress_rsvd(a, 1e-10, 16, s)
    assert [b.shape for b in s.log] == [(300, 26)]
    assert lr.rank == int(np.sum(svals > 1e-10))
    np.testing.assert_allclose(lr.U.T @ lr.U, np.eye(lr.rank), atol=1e-12)
    assert np.linalg.norm(a - lr.U @ lr.B @ lr.V.T, 2) < 1e-9


def test_fixed_rank_power_iterations_improve_slow_decay():
    svals = 1.0 / np.arange(1, 61) ** 1.0
    a = _planted(500, 500, svals, seed=9)
    errs = []
    for q in (0, 2):
        o = lowrank.Options(rank=10, fixed_rank=True, oversample=5, power_iters=q, exact_below=0)
        lr = lowrank.compress(a, o, rng.RandomStream(4))
        assert lr.rank == 10
        errs.append(np.linalg.norm(a - lr.U @ lr.B @ lr.V.T, 2))
    assert errs[1] < errs[0]
    assert errs[1] < 1.5 * svals[10]


# -------------------------------------------------------------- HODLR ----
def test_hodlr_oracle_against_dense():
    n = 1024
    a = kernels.generate_matrix(kernels.CAUCHY, n)
    s = rng.RandomStream(7, record=True)
    xt = s.random(n, 1)
    b = a @ xt
    h = hodlr.HODLRMatrix(1e-8)
    h.build_from_dense(a, 128, s)
    assert hodlr.rel_error(b, h.multiply(xt)) < 1e-9
    # matvec recursion == dense reconstruction
    np.testing.assert_allclose(h.multiply(xt), h.to_dense() @ xt, rtol=1e-13, atol=0)
    x = h.solve(b)
    assert hodlr.rel_error(h.solve_dense(b), x) < 1e-11
    assert hodlr.rel_error(np.linalg.solve(a, b), x) < 1e-6
    # draw order = post-order A12, A21 (hodlr_routines.h:43-53)
    sched = hodlr.omega_schedule(n, 128, h._options())
    assert [blk.shape for blk in s.log[1:]] == [(nn, kk) for nn, kk in sched]


def test_hodlr_trivial_tree_is_dense_lu():
    a = kernels.gaussian_points(np.linspace(0, 1, 50), 0.2, 1.0)
    h = hodlr.HODLRMatrix(1e-8)
    h.build_from_dense(a, 64, rng.RandomStream(0))
    b = np.arange(50.0).reshape(-1, 1)
    np.testing.assert_allclose(h.solve(b), np.linalg.solve(a, b), rtol=1e-12)


def test_hodlr_oracle_catches_planted_mistake():
    n = 512
    a = kernels.generate_matrix(kernels.CAUCHY, n)
    h = hodlr.HODLRMatrix(1e-8)
    h.build_from_dense(a, 64, rng.RandomStream(3))
    b = np.ones((n, 1))
    good = h.solve(b)
    # swap the coupling matrices of the root: the dense checks must notice
    h.root.B12, h.root.B21 = h.root.B21 * 1.01, h.root.B12
    bad = h.solve(b)
    assert hodlr.rel_error(good, bad) > 1e-6
    assert hodlr.rel_error(h.to_dense() @ bad, b) < 1e-10  # still self-consistent
    assert np.linalg.norm(h.to_dense() - a) / np.linalg.norm(a) > 1e-6


def test_hodlr_ragged_tree():
    n = 777
    x = kernels.sorted_uniform_points(n, 1)
    a = kernels.gaussian_points(x, 0.1, 1.0)
    h = hodlr.HODLRMatrix(1e-10)
    h.build_from_dense(a, 60, rng.RandomStream(5))
    b = np.random.default_rng(0).standard_normal((n, 3))
    np.testing.assert_allclose(h.solve(b), h.solve_dense(b), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- HSS ----
def test_hss_oracle_against_dense():
    n = 512
    a = kernels.gaussian_points(np.linspace(0, 1, n), 0.1, 1.0)
    s = rng.RandomStream(3)
    xt = s.random(n, 1)
    b = a @ xt
    h = hss.HSSMatrix(1e-9)
    h.build_from_dense(a, 64, s)
    assert hodlr.rel_error(b, h.multiply(xt)) < 1e-8
    np.testing.assert_allclose(h.multiply(xt), h.to_dense() @ xt, rtol=1e-12)
    x = h.solve(b)
    assert hodlr.rel_error(h.solve_dense(b), x) < 1e-10
    assert hodlr.rel_error(xt, x) < 1e-6
    # identity transfers: root rank = sum of leaf ranks (hss_routines.h:82-83)
    assert h.root.rank_U == sum(nd.rank_U for nd in h.nodes if nd.is_leaf)
