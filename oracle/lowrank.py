"""Low-rank compression — oracle only.

``compress_svd`` restates include/compression.h:18-64 and ``compress_rsvd``
restates include/compression.h:67-99.  ``compress`` generalises the latter
with the options the engine exposes (fixed rank, oversampling, power
iterations, exact-SVD threshold); with ``Options.reference(...)`` it is
exactly ``compress_rsvd``.

LAPACK: dgesvd('S','S') through scipy (the routine named at
compression.h:33,38).
"""
from dataclasses import dataclass

import numpy as np
import scipy.linalg as sla


@dataclass
class LowRank:
    U: np.ndarray  # m x r, orthonormal columns
    V: np.ndarray  # n x r, orthonormal columns
    B: np.ndarray  # r x r, diag(S)
    rank: int


def _gesvd(a):
    return sla.svd(a, full_matrices=False, lapack_driver="gesvd", check_finite=False)


def compress_svd(a, tol):
    """compression.h:18-64: full SVD, rank = #{S_i > tol*S_0} (>= 1)."""
    m, n = a.shape
    if m == 0 or n == 0:
        z = np.zeros((0, 0))
        return LowRank(z, z, z, 0)
    u, s, vt = _gesvd(np.asfortranarray(a))
    rank = 0
    if s[0] > 1e-15:
        rank = int(np.count_nonzero(s > tol * s[0]))
    if rank == 0:
        rank = 1
    return LowRank(np.asfortranarray(u[:, :rank]), np.asfortranarray(vt[:rank, :].T),
                   np.asfortranarray(np.diag(s[:rank])), rank)


@dataclass
class Options:
    tol: float = 1e-6          # relative truncation S_i > tol*S_0 (tol mode)
    rank: int = 64             # rsvd_rank (tol mode) or the fixed rank
    fixed_rank: bool = False   # False: reference tolerance rule
    oversample: int = 10       # k = rank + oversample  (compression.h:74)
    power_iters: int = 0       # q (the reference has none)
    exact_below: int = 256     # exact SVD if m or n < this (compression.h:72)

    @staticmethod
    def reference(tol, rsvd_rank=64):
        return Options(tol=tol, rank=rsvd_rank)


def _orth(y):
    """Left singular vectors of Y, all columns (compress_svd(Y, -1))."""
    return compress_svd(y, -1.0).U


def uses_rsvd(m, n, opts):
    """True when ``compress`` draws an Omega for an m x n block."""
    if m < opts.exact_below or n < opts.exact_below:
        return False
    k = min(opts.rank + opts.oversample, m, n)
    if opts.fixed_rank and k >= min(m, n):
        return False
    return True


def sketch_width(m, n, opts):
    return min(opts.rank + opts.oversample, m, n)


def _truncate_fixed(a_svd_u, s, vt, r):
    r = min(r, len(s))
    return a_svd_u[:, :r], vt[:r, :].T, s[:r], r


def compress(a, opts, rng):
    """Compress one block.  ``rng.random(n, k)`` supplies Omega."""
    m, n = a.shape
    if not uses_rsvd(m, n, opts):
        if not opts.fixed_rank:
            return compress_svd(a, opts.tol)
        u, s, vt = _gesvd(np.asfortranarray(a))
        u, v, s, r = _truncate_fixed(u, s, vt, opts.rank)
        return LowRank(np.asfortranarray(u), np.asfortranarray(v), np.diag(s), r)
    k = sketch_width(m, n, opts)
    omega = rng.random(n, k)
    y = a @ omega
    for _ in range(opts.power_iters):
        q = _orth(y)
        z = a.T @ q
        w = _orth(z)
        y = a @ w
    q = _orth(y)
    b_small = q.T @ a
    if not opts.fixed_rank:
        svd_b = compress_svd(b_small, opts.tol)
        return LowRank(np.asfortranarray(q @ svd_b.U), svd_b.V, svd_b.B, svd_b.rank)
    u, s, vt = _gesvd(np.asfortranarray(b_small))
    u, v, s, r = _truncate_fixed(u, s, vt, opts.rank)
    return LowRank(np.asfortranarray(q @ u), np.asfortranarray(v), np.diag(s), r)


def compress_rsvd(a, tol, rank_guess, rng):
    """compression.h:67-99 (q = 0, exact SVD below 256, k = rank_guess+10)."""
    return compress(a, Options.reference(tol, rank_guess), rng)
