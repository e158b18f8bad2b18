"""The drop-in headers (include/hmat_amd.hpp, include/hmat_amd_dropin.hpp),
CPU only: a driver written with the reference's names and call sequence
(src/main_hodlr.cpp:48-104, src/main_hss.cpp:51-109) compiles and links
against them; the class templates instantiate with a type shaped like the
reference's Matrix (include/matrix.h:28-41: public rows_/cols_/data_,
(rows, cols) constructor, rows()/cols()/data()); and the host-side pieces
(dense Matrix, Kernels:: generators) compute what the reference's do,
checked against closed forms and numpy."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "hmat-omp_amd", "lib")

DRIVER = r'''
#include <iostream>
#include <omp.h>
#include "hmat_amd_dropin.hpp"
double calc_rel_error(const Matrix& ref, const Matrix& res) {
  double num = 0, den = 0;
  for (int i = 0; i < ref.rows() * ref.cols(); ++i) {
    double d = ref.data()[i] - res.data()[i];
    num += d * d; den += ref.data()[i] * ref.data()[i];
  }
  return den == 0 ? 0.0 : std::sqrt(num) / std::sqrt(den);
}
int main() {
  omp_set_max_active_levels(4);
  int N = 4096, leaf_size = 512; double tol = 1e-8;
  Matrix A = Kernels::generate_matrix(Kernels::CAUCHY, N);
  Matrix x_true = Matrix::Random(N, 1);
  Matrix b = A * x_true;
  Matrix A_dense = A; Matrix x_blas = b; A_dense.solve_lu(x_blas);
  HODLRMatrix hodlr_serial(tol);
  hodlr_serial.build_from_dense(A, leaf_size);
  Matrix y_serial = hodlr_serial.multiply(x_true);
  Matrix x_serial = hodlr_serial.solve(b);
  HODLRMatrixOMP hodlr_omp(tol);
  hodlr_omp.rsvd_rank = 48;
  hodlr_omp.build_from_dense(A, leaf_size);
  Matrix x_omp = hodlr_omp.solve(b);
  Matrix G = Kernels::generate_gaussian(2048, 0.5, 1e-6);
  HSSMatrix hss(1e-9); hss.build_from_dense(G, 256);
  HSSMatrixOMP hss_omp(1e-9); hss_omp.build_from_dense(G, 256);
  Matrix xs = hss.solve(G * x_true.block(0, 0, 2048, 1)), xo = hss_omp.solve(G * x_true.block(0, 0, 2048, 1));
  std::cout << calc_rel_error(b, y_serial) << " " << calc_rel_error(x_true, x_omp) << " " << omp_get_max_threads()
            << " " << xs.rows() + xo.rows() << std::endl;
  return 0;
}
'''

REF_SHAPED = r'''
#include <vector>
#include "hmat_amd.hpp"
// a dense type with the reference Matrix's layout and accessors
class RefMatrix {
 public:
  int rows_, cols_;
  std::vector<double> data_;
  RefMatrix() : rows_(0), cols_(0) {}
  RefMatrix(int r, int c) : rows_(r), cols_(c), data_(r * c, 0.0) {}
  int rows() const { return rows_; }
  int cols() const { return cols_; }
  double* data() { return data_.data(); }
  const double* data() const { return data_.data(); }
};
template class hmat_amd::HODLRMatrixOMPT<RefMatrix>;
template class hmat_amd::HSSMatrixOMPT<RefMatrix>;
template class hmat_amd::HSSMatrixT<RefMatrix>;
using HODLRMatrixOMP = hmat_amd::HODLRMatrixOMPT<RefMatrix>;
using HODLRMatrix = hmat_amd::HODLRMatrixT<RefMatrix>;
RefMatrix use(const RefMatrix& A, const RefMatrix& b) {
  HODLRMatrixOMP h(1e-8);
  h.build_from_dense(A, 256);
  HODLRMatrix s;
  s.build_from_dense(A, 256);
  return h.solve(s.multiply(b));
}
'''

HOST = r'''
#include <cstdio>
#include "hmat_amd_dropin.hpp"
int main(int argc, char** argv) {
  const int n = 300;
  // generators
  Matrix C = Kernels::generate_matrix(Kernels::CAUCHY, n), G = Kernels::generate_matrix(Kernels::GAUSSIAN, n),
         K = Kernels::generate_matrix(Kernels::COULOMB, n);
  std::FILE* f = std::fopen(argv[1], "wb");
  std::fwrite(C.data(), sizeof(double), C.data_.size(), f);
  std::fwrite(G.data(), sizeof(double), G.data_.size(), f);
  std::fwrite(K.data(), sizeof(double), K.data_.size(), f);
  // dense algebra on a random system
  Matrix A = Matrix::Random(n, n) + Matrix::Identity(n, n) * Matrix::Identity(n, n);
  for (int i = 0; i < n; ++i) A(i, i) += 2.0;
  Matrix B = Matrix::Random(n, 3);
  Matrix X = B, F = A;
  F.solve_lu(X);
  Matrix Ai = A.inverse();
  Matrix T = A.transpose();
  Matrix S = A.block(10, 20, 30, 40);
  Matrix Z = Matrix::Zero(5, 5);
  Z.set_block(3, 3, Matrix::Identity(4, 4));  // clipped to the 2 x 2 corner
  for (const Matrix* m : {&A, &B, &X, &Ai, &T, &S, &Z}) std::fwrite(m->data(), sizeof(double), m->data_.size(), f);
  Matrix AB = A * B, D = A * X - B;
  std::fwrite(AB.data(), sizeof(double), AB.data_.size(), f);
  std::fwrite(D.data(), sizeof(double), D.data_.size(), f);
  std::fclose(f);
  return 0;
}
'''


def _gxx(tmp_path, name, src, link, extra=()):
    c = tmp_path / f"{name}.cpp"
    c.write_text(src)
    out = tmp_path / name
    cmd = ["g++", "-std=c++17", "-O2", "-fopenmp", "-Wall", "-I", INC, str(c)]
    if link:
        cmd += ["-L", LIBDIR, "-lhmat_amd", f"-Wl,-rpath,{LIBDIR}", "-o", str(out)]
    else:
        cmd += ["-c", "-o", str(out) + ".o"]
    cmd += list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return out


def test_reference_driver_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libhmat_amd.so")):
        pytest.skip("library not built")
    _gxx(tmp_path, "driver", DRIVER, link=True)


def test_templates_instantiate_with_reference_shaped_matrix(tmp_path):
    _gxx(tmp_path, "refshaped", REF_SHAPED, link=False)


def test_host_matrix_and_generators(tmp_path):
    exe = _gxx(tmp_path, "host", HOST, link=False, extra=())
    # link the host check without the library (it calls no engine function)
    obj = str(exe) + ".o"
    r = subprocess.run(["g++", "-fopenmp", obj, "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "host.bin"
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v = np.fromfile(out)
    n = 300
    pos = 0

    def take(rows, cols):
        nonlocal pos
        m = v[pos:pos + rows * cols].reshape(cols, rows).T
        pos += rows * cols
        return m

    C, G, K = take(n, n), take(n, n), take(n, n)
    x = np.arange(n) / (n - 1)                       # kernel.h:18-26
    d = x[:, None] - x[None, :]
    assert np.allclose(C, 1.0 / (x[:, None] + x[None, :] + 1.0) + np.eye(n), rtol=1e-15, atol=0)
    assert np.allclose(G, np.exp(-(d * d) / (0.1 * 0.1)) + 1e-6 * np.eye(n), rtol=1e-15, atol=0)
    with np.errstate(divide="ignore"):
        k = np.where(np.eye(n, dtype=bool), 0.0, 1.0 / np.abs(d)) + 10.0 * n * np.eye(n)
    assert np.allclose(K, k, rtol=1e-15, atol=0)
    A, B, X, Ai, T, S, Z = take(n, n), take(n, 3), take(n, 3), take(n, n), take(n, n), take(30, 40), take(5, 5)
    AB, D = take(n, 3), take(n, 3)
    assert np.all((B >= 0) & (B < 1))                # Matrix::Random, U[0,1)
    kappa = np.linalg.cond(A)
    assert np.linalg.norm(A @ X - B) <= 1e-14 * np.linalg.norm(A) * np.linalg.norm(X)
    assert np.linalg.norm(X - np.linalg.solve(A, B)) <= 1e-15 * kappa * np.linalg.norm(X)
    assert np.linalg.norm(Ai - np.linalg.inv(A)) <= 1e-15 * kappa * np.linalg.norm(Ai)
    assert np.array_equal(T, A.T) and np.array_equal(S, A[10:40, 20:60])
    ez = np.zeros((5, 5)); ez[3, 3] = ez[4, 4] = 1.0
    assert np.array_equal(Z, ez)
    assert np.linalg.norm(AB - A @ B) <= 1e-14 * np.linalg.norm(AB)
    assert np.linalg.norm(D) <= 1e-12 * np.linalg.norm(B)  # A (A^{-1} B) - B
