"""Batches of independent Gaussian systems (a HODLR forest).

A_s = exp(-(p_i - p_j)^2 / (2 sigma_s^2)) + lam delta_ij on sorted U[0,1]
points of system s, lam = 1e-2, leaf 64, fixed rank 24, per-system RHS
counts as zero-padded columns.  Per-system widths enter as points scaled by
1/sigma_s with a unit-width kernel.  Bars: the operator is exactly block
diagonal (cross-system blocks are 0), each diagonal block approximates its
system's kernel matrix within the rank-24 compression error, solve = dense
LAPACK on each block to 1e-9 (exact algebra), residual against A_H <= 1e-10
at the full batch size.
"""
import numpy as np
import pytest
import torch

import hmat_amd
from oracle import hodlr as oh, kernels, lowrank, rng as rng_o

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _batch(nsys, nlo, nhi, seed):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nlo), np.log(nhi), nsys)).astype(np.int64)
    sigmas = rng.uniform(0.05, 0.3, nsys)
    nrhs = rng.integers(1, 65, nsys)
    pts = [np.sort(rng.random(n)) for n in sizes]
    return sizes, sigmas, nrhs, pts


def _opts():
    return hmat_amd.make_options(0.0, rank=24, fixed_rank=True, oversample=10, power_iters=1,
                                 exact_below=0, omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=99)


def _build(sizes, sigmas, pts, lam):
    scaled = np.concatenate([p / s for p, s in zip(pts, sigmas)])
    dp = torch.tensor(scaled, device="cuda")
    h = hmat_amd.HODLRMatrixOMP(opts=_opts())
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_GAUSSIAN, dp.data_ptr(), p0=1.0, p1=lam)
    torch.cuda.synchronize()
    h.build_kernel_batch_device(k, sizes, 64)
    return h, dp


def test_small_batch_block_diagonal_and_dense():
    lam = 1e-2
    sizes, sigmas, nrhs, pts = _batch(10, 200, 1500, 3)
    h, _ = _build(sizes, sigmas, pts, lam)
    ad = h.to_dense()
    offs = np.concatenate([[0], np.cumsum(sizes)])
    mask = np.zeros_like(ad, dtype=bool)
    for s in range(len(sizes)):
        mask[offs[s]:offs[s + 1], offs[s]:offs[s + 1]] = True
    assert np.all(ad[~mask] == 0.0)
    rmax = int(nrhs.max())
    rng = np.random.default_rng(4)
    b = np.zeros((offs[-1], rmax))
    for s in range(len(sizes)):
        b[offs[s]:offs[s + 1], :nrhs[s]] = rng.standard_normal((sizes[s], nrhs[s]))
    x = h.solve(b)
    for s in range(len(sizes)):
        sl = slice(offs[s], offs[s + 1])
        blk = ad[sl, sl]
        a_true = kernels.gaussian_points(pts[s], sigmas[s], lam)
        # the oracle's HODLR of this system with the same options (its own
        # Omega) sets the compression error of rank 24 at this sigma
        ho = oh.HODLRMatrix(opts=lowrank.Options(tol=0.0, rank=24, fixed_rank=True, oversample=10,
                                                 power_iters=1, exact_below=0))
        ho.build_from_dense(a_true, 64, rng_o.RandomStream(100 + s))
        e_gpu, e_ora = rel(blk, a_true), rel(ho.to_dense(), a_true)
        print(f"system {s}: N={sizes[s]} sigma={sigmas[s]:.3f}: block vs A {e_gpu:.1e}, "
              f"oracle vs A {e_ora:.1e}", flush=True)
        assert e_gpu <= 10 * e_ora + 1e-14
        assert rel(x[sl, :nrhs[s]], np.linalg.solve(blk, b[sl, :nrhs[s]])) < 1e-9
        assert np.all(x[sl, nrhs[s]:] == 0.0)


def test_cfg4_batch_residuals():
    """256 systems, N log-uniform in [1024, 16384], sigma ~ U[0.05, 0.3],
    lam = 1e-2, leaf 64, rank 24, RHS counts in [1, 64] (zero-padded)."""
    import time
    lam = 1e-2
    sizes, sigmas, nrhs, pts = _batch(256, 1024, 16384, 2024)
    n = int(sizes.sum())
    rmax = int(nrhs.max())
    rng = np.random.default_rng(5)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    bh = np.zeros((rmax, n))  # column-major storage of an n x rmax matrix
    for s in range(len(sizes)):
        bh[:nrhs[s], offs[s]:offs[s + 1]] = rng.standard_normal((nrhs[s], sizes[s]))
    db = torch.tensor(bh, device="cuda")
    dx = torch.empty_like(db)
    dy = torch.empty_like(db)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    h, _ = _build(sizes, sigmas, pts, lam)
    # kappa(A_s) reaches ~1e6 (sigma 0.3, N 16384, lam 1e-2): FP64
    # refinement takes every column's residual below 1e-11 (the per-system
    # bar below is 1e-10)
    h.set_refinement(1e-11, 10)
    t1 = time.perf_counter()
    h.solve_device(db.data_ptr(), n, rmax, dx.data_ptr(), n)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    print(f"cfg4: {len(sizes)} systems, sum N = {n}, {int(nrhs.sum())} RHS: build {t1 - t0:.3f}s "
          f"factor+solve {t2 - t1:.3f}s, refinement {h.last_refinement()}", flush=True)
    h.multiply_device(dx.data_ptr(), n, rmax, dy.data_ptr(), n)
    torch.cuda.synchronize()
    r = (db - dy).cpu().numpy()
    bb = bh
    for s in range(len(sizes)):
        sl = slice(offs[s], offs[s + 1])
        for j in range(nrhs[s]):
            assert np.linalg.norm(r[j, sl]) <= 1e-10 * np.linalg.norm(bb[j, sl])


def test_batch_save_load(tmp_path):
    sizes, sigmas, nrhs, pts = _batch(5, 300, 900, 8)
    h, _ = _build(sizes, sigmas, pts, 1e-2)
    f = tmp_path / "batch.hodlr"
    h.save(f)
    g = hmat_amd.HODLRMatrixOMP(1e-3)
    g.load(f)
    assert np.array_equal(g.to_dense(), h.to_dense())


def test_autotuned_batch_build_matches_default():
    """With the tile autotuner on, the batch operator equals the default
    build up to rounding (the tuned tilings only regroup the same sums)."""
    sizes, sigmas, nrhs, pts = _batch(24, 1024, 4096, 12)
    hd, _ = _build(sizes, sigmas, pts, 1e-2)
    a0 = hd.to_dense()
    hmat_amd.autotune_enable(True)
    try:
        before = hmat_amd.autotune_entries()
        h1, _ = _build(sizes, sigmas, pts, 1e-2)   # tunes new shape classes
        assert hmat_amd.autotune_entries() > before
        h2, _ = _build(sizes, sigmas, pts, 1e-2)   # reuses the table
        a1, a2 = h1.to_dense(), h2.to_dense()
    finally:
        hmat_amd.autotune_enable(False)
    assert rel(a1, a0) < 1e-12
    assert np.array_equal(a1, a2)
