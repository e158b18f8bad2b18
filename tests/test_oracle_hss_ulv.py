"""CPU checks of the oracle's ULV factorisation (oracle/hss_ulv.py), pinned
by what the mathematics fixes rather than by a second copy of its formulas:

  * dense LAPACK on the operator's own reconstruction: both solve the same
    system, so the solutions agree to kappa * eps, and the ULV backward
    error stays within a small multiple of LAPACK's (an orthogonal
    elimination is backward stable), including a Cauchy system with kappa
    about 3e9;
  * an exactly HSS matrix (block rows of exact low rank): the solve equals
    the solve with A itself;
  * the transforms: Qt is orthogonal and maps the basis to [I; 0]; the
    eliminated rows carry no coupling (U_perp^T U = 0); C[r:]^T = H_P [Rq; 0];
  * a tree of one leaf is a plain QR solve;
  * planted mistake: leaving the sibling's known part out of the parent's
    right-hand side moves the solution by 1e-4 or more.
"""
import numpy as np
import pytest

from oracle.hss_sampled import HSSSampled
from oracle.hss_ulv import HSSULV


def system(kind, n, seed=1):
    rng = np.random.default_rng(seed)
    x = np.sort(rng.random(n))
    if kind == "cauchy":
        return 1.0 / (x[:, None] - (x + 0.5 / n)[None, :])
    return np.exp(-(x[:, None] - x[None, :]) ** 2 / (2 * 0.05 ** 2)) + np.eye(n)


def build(a, leaf, rank, seed=2):
    n = a.shape[0]
    s = (rank or 40) + 10
    rng = np.random.default_rng(seed)
    return HSSSampled(rank=rank, tol=1e-12).build_from_dense(a, leaf, rng.standard_normal((n, s)),
                                                             rng.standard_normal((n, s)))


def backward(a, x, b):
    return np.linalg.norm(a @ x - b) / (np.linalg.norm(a, 2) * np.linalg.norm(x))


@pytest.mark.parametrize("kind,n,leaf,rank", [("cauchy", 512, 64, 0), ("gauss", 1000, 100, 20),
                                              ("gauss", 777, 50, 0)])
def test_against_lapack(kind, n, leaf, rank):
    h = build(system(kind, n), leaf, rank)
    ah = h.to_dense()
    b = np.random.default_rng(3).standard_normal((n, 3))
    x = HSSULV(h).solve(b)
    xl = np.linalg.solve(ah, b)
    kappa = np.linalg.cond(ah)
    assert backward(ah, x, b) < 50 * backward(ah, xl, b) + 1e-16
    assert np.linalg.norm(x - xl) / np.linalg.norm(xl) < 1e-14 * kappa


def test_exact_hss_matrix():
    n = 640
    g = np.random.default_rng(4)
    u, v = g.standard_normal((n, 4)), g.standard_normal((n, 4))
    a = u @ v.T + n * np.eye(n)
    h = build(a, 64, 0)
    b = g.standard_normal((n, 2))
    x = HSSULV(h).solve(b)
    assert np.linalg.norm(x - np.linalg.solve(a, b)) / np.linalg.norm(x) < 1e-13


def test_transforms():
    h = build(system("gauss", 512), 64, 16)
    f = HSSULV(h).f
    for i, nd in enumerate(h.nodes):
        fi = f[i]
        n, r, m = fi["n"], fi["r"], fi["m"]
        assert np.linalg.norm(fi["Qt"] @ fi["Qt"].T - np.eye(n)) < 1e-13
        assert np.linalg.norm(fi["Hp"] @ fi["Hp"].T - np.eye(n)) < 1e-13
        if i and nd.left < 0:
            qu = fi["Qt"] @ h.U[i]
            assert np.linalg.norm(qu[:r] - np.eye(r)) < 1e-12
            assert np.linalg.norm(qu[r:]) < 1e-12
        if m:
            assert np.allclose(np.tril(fi["Rq"], -1), 0.0)
            c = fi["Qt"] @ (h.D[i] if nd.left < 0 else np.eye(n))
            if nd.left < 0:
                rq = np.zeros((n, m))
                rq[:m] = fi["Rq"]
                assert np.linalg.norm(c[r:].T - fi["Hp"] @ rq) < 1e-12 * np.linalg.norm(c)


def test_single_leaf_is_qr_solve():
    a = system("gauss", 50)
    h = build(a, 64, 0)
    b = np.random.default_rng(5).standard_normal((50, 2))
    assert np.linalg.norm(HSSULV(h).solve(b) - np.linalg.solve(a, b)) < 1e-12


def test_planted_mistake_is_caught():
    h = build(system("gauss", 1000), 100, 20)
    b = np.random.default_rng(6).standard_normal((1000, 2))
    good, bad = HSSULV(h).solve(b), HSSULV(h, skip_sibling=True).solve(b)
    xl = np.linalg.solve(h.to_dense(), b)
    assert np.linalg.norm(good - xl) / np.linalg.norm(xl) < 1e-12
    assert np.linalg.norm(bad - xl) / np.linalg.norm(xl) > 1e-4
