"""Kernel-matrix generators — oracle only.

Reference generators (include/kernel.h):
  * grid points x_i = i/(N-1)                                   kernel.h:18-26
  * Cauchy   A_ij = 1/(x_i + x_j + 1) (+ shift on the diagonal)  kernel.h:30-48
  * Gaussian A_ij = exp(-(x_i-x_j)^2 / h^2) + shift * delta_ij   kernel.h:52-65
  * Coulomb  A_ij = 1/|x_i-x_j|, 0 on the diagonal, + shift      kernel.h:69-89
  * presets  Cauchy shift 1, Gaussian h=0.1 shift 1e-6,
             Coulomb shift 10N                                  kernel.h:92-106

Point-based kernels of the benchmark configurations (no reference
counterpart; closed forms):
  * gaussian_points: A_ij = exp(-(x_i-x_j)^2 / (2 sigma^2)) + lam * delta_ij
  * cauchy_points:   A_ij = 1/(x_i - y_j), x_i = i/N, y_j = (j+1/2)/N
"""
import numpy as np

CAUCHY, GAUSSIAN, COULOMB = 0, 1, 2


def generate_grid(n):
    return np.arange(n, dtype=np.float64) / float(n - 1)


def generate_cauchy(n, diagonal_shift=0.0):
    x = generate_grid(n)
    a = 1.0 / (x[:, None] + x[None, :] + 1.0)
    if diagonal_shift != 0.0:
        a[np.diag_indices(n)] += diagonal_shift
    return np.asfortranarray(a)


def generate_gaussian(n, bandwidth, diagonal_shift=1e-12):
    x = generate_grid(n)
    d = x[:, None] - x[None, :]
    a = np.exp(-(d * d) / (bandwidth * bandwidth))
    a[np.diag_indices(n)] += diagonal_shift
    return np.asfortranarray(a)


def generate_coulomb(n, diagonal_shift=0.0):
    x = generate_grid(n)
    d = np.abs(x[:, None] - x[None, :])
    with np.errstate(divide="ignore"):
        a = 1.0 / d
    a[np.diag_indices(n)] = 0.0
    if diagonal_shift != 0.0:
        a[np.diag_indices(n)] += diagonal_shift
    return np.asfortranarray(a)


def generate_matrix(kind, n):
    if kind == CAUCHY:
        return generate_cauchy(n, 1.0)
    if kind == GAUSSIAN:
        return generate_gaussian(n, 0.1, 1e-6)
    if kind == COULOMB:
        return generate_coulomb(n, 10.0 * n)
    return np.asfortranarray(np.eye(n))


def sorted_uniform_points(n, seed):
    """N points sorted from U[0,1) (seeded numpy PCG64 stream)."""
    rs = np.random.default_rng(seed)
    return np.sort(rs.random(n))


def gaussian_points(x, sigma, lam, rows=None, cols=None):
    """exp(-(x_i-x_j)^2/(2 sigma^2)) + lam*delta_ij on the given index sets."""
    rows = np.arange(len(x)) if rows is None else np.asarray(rows)
    cols = np.arange(len(x)) if cols is None else np.asarray(cols)
    d = x[rows][:, None] - x[cols][None, :]
    a = np.exp(-(d * d) / (2.0 * sigma * sigma))
    a = a + lam * (rows[:, None] == cols[None, :])
    return np.asfortranarray(a)


def cauchy_grid_points(n):
    """x_i = i/N and y_j = (j+1/2)/N (both exact in binary for N = 2^p)."""
    i = np.arange(n, dtype=np.float64)
    return i / n, (i + 0.5) / n


def cauchy_points(x, y, rows=None, cols=None):
    rows = np.arange(len(x)) if rows is None else np.asarray(rows)
    cols = np.arange(len(y)) if cols is None else np.asarray(cols)
    return np.asfortranarray(1.0 / (x[rows][:, None] - y[cols][None, :]))
