// A driver in the form of the reference drivers (src/main_hodlr.cpp,
// src/main_hss.cpp), compiled against include/hmat_amd_dropin.hpp instead of
// the reference headers: the same call sequence on the same names (Matrix,
// Kernels::, HODLRMatrix / HODLRMatrixOMP, HSSMatrix / HSSMatrixOMP), the
// dense baseline on the host, the hierarchical solvers on the MI355X.
//
//   drop_in_driver hodlr [N leaf tol]   defaults 4096 512 1e-8, the drivers'
//                                       Cauchy matrix (main_hodlr.cpp:38-49)
//   drop_in_driver hss [N leaf tol]     defaults 2048 256 1e-9, Gaussian
//                                       h 0.5, shift 1e-6 (main_hss.cpp:38-52)
//
// Prints the reference's lines (build / GEMV / solve times, relative errors
// of the product against b and of the solution against x_true) and, for
// each solver, the residual ||b - A x|| / ||b||, which the reference does not
// print (its HSS test matrix is too ill-conditioned for the error against
// x_true to mean anything).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "hmat_amd_dropin.hpp"

namespace {

struct Stopwatch {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double lap() {
    const auto t1 = std::chrono::steady_clock::now();
    const double s = std::chrono::duration<double>(t1 - t0).count();
    t0 = t1;
    return s;
  }
};

// Frobenius relative error, as calc_rel_error (main_hodlr.cpp:26-35)
double rel_error(const Matrix& ref, const Matrix& res) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < ref.data_.size(); ++i) {
    const double d = ref.data_[i] - res.data_[i];
    num += d * d;
    den += ref.data_[i] * ref.data_[i];
  }
  return den == 0.0 ? 0.0 : std::sqrt(num) / std::sqrt(den);
}

template <class Solver>
void run(const char* title, Solver& h, const Matrix& A, int leaf, const Matrix& x_true, const Matrix& b,
         double* t_build, double* t_solve) {
  std::printf("\n--- %s ---\n", title);
  Stopwatch sw;
  h.build_from_dense(A, leaf);
  *t_build = sw.lap();
  std::printf(" Build: %.4fs\n", *t_build);
  Matrix y = h.multiply(x_true);
  const double t_mv = sw.lap();
  std::printf(" GEMV:  %.4fs | Err: %.6e\n", t_mv, rel_error(b, y));
  Matrix x = h.solve(b);
  *t_solve = sw.lap();
  std::printf(" Solve: %.4fs | Err: %.6e\n", *t_solve, rel_error(x_true, x));
  std::printf(" Resid: %.6e\n", rel_error(b, A * x));
}

}  // namespace

int main(int argc, char** argv) {
  const std::string kind = argc > 1 ? argv[1] : "hodlr";
  const bool hss = kind == "hss";
  if (!hss && kind != "hodlr") {
    std::fprintf(stderr, "usage: drop_in_driver hodlr|hss [N leaf tol]\n");
    return 2;
  }
  const int N = argc > 2 ? std::atoi(argv[2]) : (hss ? 2048 : 4096);
  const int leaf = argc > 3 ? std::atoi(argv[3]) : (hss ? 256 : 512);
  const double tol = argc > 4 ? std::atof(argv[4]) : (hss ? 1e-9 : 1e-8);
  if (hmat_device_count() < 1) {
    std::fprintf(stderr, "no HIP device\n");
    return 2;
  }
  std::printf("%s benchmark on the MI355X engine: N=%d leaf=%d tol=%g\n", hss ? "HSS" : "HODLR", N, leaf, tol);
  Matrix A = hss ? Kernels::generate_gaussian(N, 0.5, 1e-6) : Kernels::generate_matrix(Kernels::CAUCHY, N);
  Matrix x_true = Matrix::Random(N, 1);
  Matrix b = A * x_true;

  // the dense baseline (main_hodlr.cpp:53-65), on the host
  Stopwatch sw;
  Matrix b_dense = A * x_true;
  const double t_dense_mv = sw.lap();
  Matrix A_dense = A;
  Matrix x_dense = b;
  A_dense.solve_lu(x_dense);
  const double t_dense_solve = sw.lap();
  std::printf("\n--- Dense baseline (host) ---\n GEMV: %.4fs | Solve: %.4fs | Err: %.6e\n", t_dense_mv, t_dense_solve,
              rel_error(x_true, x_dense));
  double tb_s = 0, ts_s = 0, tb_p = 0, ts_p = 0;
  try {
    if (hss) {
      HSSMatrix serial(tol);
      run("HSSMatrix (MI355X)", serial, A, leaf, x_true, b, &tb_s, &ts_s);
      HSSMatrixOMP omp(tol);
      run("HSSMatrixOMP (MI355X)", omp, A, leaf, x_true, b, &tb_p, &ts_p);
    } else {
      HODLRMatrix serial(tol);
      run("HODLRMatrix (MI355X)", serial, A, leaf, x_true, b, &tb_s, &ts_s);
      HODLRMatrixOMP omp(tol);
      run("HODLRMatrixOMP (MI355X)", omp, A, leaf, x_true, b, &tb_p, &ts_p);
    }
  } catch (const hmat_amd::Error& e) {
    std::fprintf(stderr, "hmat_amd error %d: %s\n", (int)e.status, e.what());
    return 1;
  }
  std::printf("\n>>> Speedup <<<\nSolve vs dense (host LU): %.2fx\n", t_dense_solve / ts_p);
  return 0;
}
