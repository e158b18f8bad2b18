// hmat_amd_dropin.hpp — everything a driver written against chenxinye/hmat-omp
// uses, on top of the MI355X engine (hmat_amd.hpp / hmat_amd.h).
//
// A reference driver (src/main_hodlr.cpp, src/main_hss.cpp) replaces its three
// includes
//     #include "../include/hodlr_routines.h"       (or hss_routines.h)
//     #include "../include/hodlr_routines_omp.h"   (or hss_routines_omp.h)
//     #include "../include/kernel.h"
// by
//     #include "hmat_amd_dropin.hpp"
// and compiles unchanged: the names below are the reference's.
//
//   Matrix           dense column-major FP64 matrix with the interface of
//                    include/matrix.h:28-150 (constructors, (i, j), rows/cols/
//                    data, setZero, Zero/Random/Identity, block, set_block,
//                    transpose, *, +, -, solve_lu, inverse).  Dense products
//                    and LU run on the host (they are the drivers' dense
//                    baseline, main_hodlr.cpp:53-65): cache-blocked loops,
//                    OpenMP-parallel when compiled with -fopenmp.
//   Kernels::        the generators of include/kernel.h:10-106 (grid, Cauchy,
//                    Gaussian, Coulomb, generate_matrix with the same presets).
//   HODLRMatrix      serial HODLR class (include/hodlr_routines.h:8-176) and
//   HODLRMatrixOMP   its OpenMP twin (include/hodlr_routines_omp.h:9-241):
//                    both are the GPU engine (same algebra, one device).
//   HSSMatrix        serial HSS class (include/hss_routines.h:10-307; solve
//                    takes every column, hss_routines.h:197-199) and
//   HSSMatrixOMP     its OpenMP twin (include/hss_routines_omp.h:11-375; solve
//                    takes column 0).
//
// Define HMAT_AMD_NO_GLOBAL_NAMES to keep only the hmat_amd:: names.
#ifndef HMAT_AMD_DROPIN_HPP
#define HMAT_AMD_DROPIN_HPP

#include <algorithm>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "hmat_amd.hpp"

namespace hmat_amd {

// Dense FP64 matrix, element (i, j) at data_[j * rows_ + i] (matrix.h:36-37).
// Public rows_/cols_/data_ as in the reference; sizes are int like the
// reference's, with products formed in 64 bits.
class Matrix {
 public:
  int rows_ = 0, cols_ = 0;
  std::vector<double> data_;

  Matrix() = default;
  Matrix(int r, int c) : rows_(r), cols_(c), data_((size_t)r * (size_t)c, 0.0) {}

  double& operator()(int i, int j) { return data_[(size_t)j * rows_ + i]; }
  const double& operator()(int i, int j) const { return data_[(size_t)j * rows_ + i]; }
  int rows() const { return rows_; }
  int cols() const { return cols_; }
  double* data() { return data_.data(); }
  const double* data() const { return data_.data(); }
  void setZero() { std::fill(data_.begin(), data_.end(), 0.0); }

  static Matrix Zero(int r, int c) { return Matrix(r, c); }
  // U[0, 1) from a per-thread Mersenne twister seeded from random_device
  // (matrix.h:50-57)
  static Matrix Random(int r, int c) {
    static thread_local std::mt19937 gen(std::random_device{}());
    std::uniform_real_distribution<double> u(0.0, 1.0);
    Matrix m(r, c);
    for (double& v : m.data_) v = u(gen);
    return m;
  }
  static Matrix Identity(int r, int c) {
    Matrix m(r, c);
    for (int i = 0; i < std::min(r, c); ++i) m(i, i) = 1.0;
    return m;
  }

  Matrix block(int r0, int c0, int nr, int nc) const {
    Matrix b(nr, nc);
    for (int j = 0; j < nc; ++j)
      std::copy_n(&(*this)(r0, c0 + j), nr, &b(0, j));
    return b;
  }
  // entries falling outside this matrix are dropped (matrix.h:76-83)
  void set_block(int r0, int c0, const Matrix& b) {
    const int nr = std::min(b.rows_, rows_ - r0), nc = std::min(b.cols_, cols_ - c0);
    for (int j = 0; j < nc; ++j)
      for (int i = 0; i < nr; ++i) (*this)(r0 + i, c0 + j) = b(i, j);
  }
  Matrix transpose() const {
    Matrix t(cols_, rows_);
    for (int j = 0; j < cols_; ++j)
      for (int i = 0; i < rows_; ++i) t(j, i) = (*this)(i, j);
    return t;
  }

  Matrix operator*(const Matrix& b) const {
    if (cols_ != b.rows_) throw std::invalid_argument("Matrix::operator*: inner dimensions differ");
    Matrix c(rows_, b.cols_);
    gemm_acc(rows_, b.cols_, cols_, data(), rows_, b.data(), b.rows_, c.data(), rows_);
    return c;
  }
  Matrix operator+(const Matrix& b) const { return axpy(b, 1.0); }
  Matrix operator-(const Matrix& b) const { return axpy(b, -1.0); }

  // b <- A^{-1} b, A overwritten by its LU factors (partial pivoting), as
  // Matrix::solve_lu (matrix.h:120-131).  A zero pivot throws (the
  // reference ignores dgetrf's info).
  void solve_lu(Matrix& b) {
    if (rows_ != cols_ || rows_ != b.rows_) throw std::invalid_argument("Matrix::solve_lu: shape mismatch");
    std::vector<int> piv;
    lu_factor(piv);
    lu_apply(piv, b);
  }
  // (matrix.h:133-149)
  Matrix inverse() const {
    if (rows_ != cols_) throw std::invalid_argument("Matrix::inverse: not square");
    Matrix f = *this, inv = Identity(rows_, rows_);
    std::vector<int> piv;
    f.lu_factor(piv);
    f.lu_apply(piv, inv);
    return inv;
  }

 private:
  Matrix axpy(const Matrix& b, double s) const {
    if (rows_ != b.rows_ || cols_ != b.cols_) throw std::invalid_argument("Matrix: shape mismatch");
    Matrix c(rows_, cols_);
    for (size_t i = 0; i < data_.size(); ++i) c.data_[i] = data_[i] + s * b.data_[i];
    return c;
  }

  // C(m x n) += A(m x k) B(k x n), column-major, blocked over k and columns
  static void gemm_acc(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C,
                       int ldc, double alpha = 1.0) {
    constexpr int KB = 128;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if ((int64_t)m * n * k > (int64_t)1 << 22)
#endif
    for (int j = 0; j < n; ++j) {
      double* c = C + (size_t)j * ldc;
      for (int k0 = 0; k0 < k; k0 += KB) {
        const int k1 = std::min(k, k0 + KB);
        for (int p = k0; p < k1; ++p) {
          const double bp = alpha * B[(size_t)j * ldb + p];
          if (bp == 0.0) continue;
          const double* a = A + (size_t)p * lda;
          for (int i = 0; i < m; ++i) c[i] += a[i] * bp;
        }
      }
    }
  }

  // Right-looking blocked LU with partial pivoting, panels of NB columns.
  void lu_factor(std::vector<int>& piv) {
    const int n = rows_;
    constexpr int NB = 64;
    piv.resize(n);
    double* a = data();
    const size_t ld = (size_t)n;
    for (int j0 = 0; j0 < n; j0 += NB) {
      const int j1 = std::min(n, j0 + NB);
      // panel: unblocked LU of columns j0..j1 over rows j0..n
      for (int j = j0; j < j1; ++j) {
        int p = j;
        double best = std::fabs(a[j * ld + j]);
        for (int i = j + 1; i < n; ++i)
          if (std::fabs(a[j * ld + i]) > best) best = std::fabs(a[j * ld + i]), p = i;
        if (best == 0.0) throw std::runtime_error("Matrix::solve_lu: zero pivot in column " + std::to_string(j));
        piv[j] = p;
        if (p != j)
          for (int c = 0; c < n; ++c) std::swap(a[c * ld + j], a[c * ld + p]);
        const double r = 1.0 / a[j * ld + j];
        for (int i = j + 1; i < n; ++i) a[j * ld + i] *= r;
        for (int c = j + 1; c < j1; ++c) {
          const double u = a[c * ld + j];
          if (u != 0.0)
            for (int i = j + 1; i < n; ++i) a[c * ld + i] -= a[j * ld + i] * u;
        }
      }
      if (j1 >= n) break;
      // U12 = L11^{-1} A12 (unit lower triangle), then A22 -= L21 U12
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
      for (int c = j1; c < n; ++c) {
        double* col = a + c * ld;
        for (int j = j0; j < j1; ++j) {
          const double u = col[j];
          if (u != 0.0)
            for (int i = j + 1; i < j1; ++i) col[i] -= a[j * ld + i] * u;
        }
        for (int j = j0; j < j1; ++j) {
          const double u = col[j];
          if (u != 0.0) {
            const double* l = a + j * ld;
            for (int i = j1; i < n; ++i) col[i] -= l[i] * u;
          }
        }
      }
    }
  }

  // b <- U^{-1} L^{-1} P b
  void lu_apply(const std::vector<int>& piv, Matrix& b) const {
    const int n = rows_;
    const double* a = data();
    const size_t ld = (size_t)n;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) if (b.cols_ > 1)
#endif
    for (int c = 0; c < b.cols_; ++c) {
      double* x = &b(0, c);
      for (int j = 0; j < n; ++j)
        if (piv[j] != j) std::swap(x[j], x[piv[j]]);
      for (int j = 0; j < n; ++j) {
        const double v = x[j];
        if (v != 0.0)
          for (int i = j + 1; i < n; ++i) x[i] -= a[j * ld + i] * v;
      }
      for (int j = n - 1; j >= 0; --j) {
        x[j] /= a[j * ld + j];
        const double v = x[j];
        if (v != 0.0)
          for (int i = 0; i < j; ++i) x[i] -= a[j * ld + i] * v;
      }
    }
  }
};

// The generators of include/kernel.h:10-106, same formulas and presets.
namespace Kernels {
enum Type { CAUCHY, GAUSSIAN, COULOMB };

// x_i = i / (N - 1) (kernel.h:18-26)
inline std::vector<double> generate_grid(int N) {
  std::vector<double> x(N);
  for (int i = 0; i < N; ++i) x[i] = (double)i / (double)(N - 1);
  return x;
}
// 1 / (x_i + x_j + 1), diagonal + shift (kernel.h:30-48)
inline Matrix generate_cauchy(int N, double diagonal_shift = 0.0) {
  const std::vector<double> x = generate_grid(N);
  Matrix A(N, N);
  for (int j = 0; j < N; ++j) {
    for (int i = 0; i < N; ++i) A(i, j) = 1.0 / (x[i] + x[j] + 1.0);
    A(j, j) += diagonal_shift;
  }
  return A;
}
// exp(-(x_i - x_j)^2 / h^2), diagonal + shift (kernel.h:52-65)
inline Matrix generate_gaussian(int N, double bandwidth, double diagonal_shift = 1e-12) {
  const std::vector<double> x = generate_grid(N);
  Matrix A(N, N);
  const double h2 = bandwidth * bandwidth;
  for (int j = 0; j < N; ++j) {
    for (int i = 0; i < N; ++i) {
      const double d = x[i] - x[j];
      A(i, j) = std::exp(-(d * d) / h2);
    }
    A(j, j) += diagonal_shift;
  }
  return A;
}
// 1 / |x_i - x_j|, 0 on the diagonal, diagonal + shift (kernel.h:69-89)
inline Matrix generate_coulomb(int N, double diagonal_shift = 0.0) {
  const std::vector<double> x = generate_grid(N);
  Matrix A(N, N);
  for (int j = 0; j < N; ++j) {
    for (int i = 0; i < N; ++i) A(i, j) = (i == j) ? 0.0 : 1.0 / std::fabs(x[i] - x[j]);
    A(j, j) += diagonal_shift;
  }
  return A;
}
// presets: Cauchy shift 1, Gaussian h 0.1 shift 1e-6, Coulomb shift 10 N
// (kernel.h:92-106)
inline Matrix generate_matrix(Type type, int N) {
  switch (type) {
    case CAUCHY: return generate_cauchy(N, 1.0);
    case GAUSSIAN: return generate_gaussian(N, 0.1, 1e-6);
    case COULOMB: return generate_coulomb(N, 10.0 * N);
  }
  return Matrix::Identity(N, N);
}
}  // namespace Kernels

}  // namespace hmat_amd

#ifndef HMAT_AMD_NO_GLOBAL_NAMES
using Matrix = hmat_amd::Matrix;
namespace Kernels = hmat_amd::Kernels;
using HODLRMatrix = hmat_amd::HODLRMatrixT<hmat_amd::Matrix>;
using HODLRMatrixOMP = hmat_amd::HODLRMatrixOMPT<hmat_amd::Matrix>;
using HSSMatrix = hmat_amd::HSSMatrixT<hmat_amd::Matrix>;
using HSSMatrixOMP = hmat_amd::HSSMatrixOMPT<hmat_amd::Matrix>;
#endif

#endif  // HMAT_AMD_DROPIN_HPP
