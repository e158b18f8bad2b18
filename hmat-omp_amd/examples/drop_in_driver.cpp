// Driver-style use of the C++ drop-in classes (include/hmat_amd.hpp): the
// flow of the reference drivers (src/main_hodlr.cpp:37-107,
// src/main_hss.cpp:51-111) — generate the Cauchy test matrix, b = A x_true,
// build, multiply, solve, and print the relative errors — with the
// reference classes replaced by hmat_amd::HODLRMatrixOMPT / HSSMatrixOMPT.
//
//   drop_in_driver [N] [leaf] [tol]        (defaults 4096 512 1e-8)
//
// The dense matrix type below has the reference Matrix's interface
// (rows(), cols(), data(), column-major, (r, c) constructor); the reference's
// own Matrix works the same way.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "hmat_amd.hpp"

struct Dense {
  int r = 0, c = 0;
  std::vector<double> v;
  Dense() = default;
  Dense(int rows, int cols) : r(rows), c(cols), v((size_t)rows * cols, 0.0) {}
  int rows() const { return r; }
  int cols() const { return c; }
  double* data() { return v.data(); }
  const double* data() const { return v.data(); }
  double& operator()(int i, int j) { return v[(size_t)j * r + i]; }
  double operator()(int i, int j) const { return v[(size_t)j * r + i]; }
};

// The drivers' test matrix: 1/(x_i + x_j + 1) on the grid x_i = i/(N-1),
// diagonal shifted by 1 (kernel.h:18-48, 92-106).
static Dense cauchy_matrix(int n) {
  Dense a(n, n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a(i, j) = 1.0 / (i / double(n - 1) + j / double(n - 1) + 1.0) + (i == j ? 1.0 : 0.0);
  return a;
}

static double rel_error(const Dense& ref, const Dense& res) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < ref.v.size(); ++i) {
    const double d = ref.v[i] - res.v[i];
    num += d * d;
    den += ref.v[i] * ref.v[i];
  }
  return den == 0.0 ? 0.0 : std::sqrt(num / den);
}

static Dense gemv(const Dense& a, const Dense& x) {
  Dense y(a.rows(), 1);
  for (int j = 0; j < a.cols(); ++j)
    for (int i = 0; i < a.rows(); ++i) y.v[i] += a(i, j) * x.v[j];
  return y;
}

template <class H>
static void run(const char* name, H& h, const Dense& A, int leaf, const Dense& x_true, const Dense& b) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  h.build_from_dense(A, leaf);
  auto t1 = clk::now();
  Dense y = h.multiply(x_true);
  auto t2 = clk::now();
  Dense x = h.solve(b);
  auto t3 = clk::now();
  auto s = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
  std::printf("--- %s ---\n Build: %.4fs\n GEMV:  %.4fs | Err: %.3e\n Solve: %.4fs | Err: %.3e\n", name,
              s(t0, t1), s(t1, t2), rel_error(b, y), s(t2, t3), rel_error(x_true, x));
}

int main(int argc, char** argv) {
  const int n = argc > 1 ? std::atoi(argv[1]) : 4096;
  const int leaf = argc > 2 ? std::atoi(argv[2]) : 512;
  const double tol = argc > 3 ? std::atof(argv[3]) : 1e-8;
  if (hmat_device_count() < 1) {
    std::fprintf(stderr, "no HIP device\n");
    return 2;
  }
  std::printf("Matrix: %dx%d (Cauchy), leaf %d, tol %g\n", n, n, leaf, tol);
  const Dense A = cauchy_matrix(n);
  Dense x_true(n, 1);
  std::mt19937 gen(12345);
  std::uniform_real_distribution<double> u(0.0, 1.0);
  for (auto& v : x_true.v) v = u(gen);
  const Dense b = gemv(A, x_true);
  try {
    hmat_amd::HODLRMatrixOMPT<Dense> hodlr(tol);
    run("HODLR (MI355X)", hodlr, A, leaf, x_true, b);
    hmat_amd::HSSMatrixOMPT<Dense> hss(tol);
    run("HSS (MI355X, reference operator)", hss, A, leaf, x_true, b);
  } catch (const hmat_amd::Error& e) {
    std::fprintf(stderr, "hmat_amd error %d: %s\n", (int)e.status, e.what());
    return 1;
  }
  return 0;
}
