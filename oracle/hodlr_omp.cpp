// hodlr_omp.cpp — CPU restatement of HODLRMatrixOMP (include/hodlr_routines_omp.h)
// with the compression of include/compression.h, for the oracle and the CPU
// baseline only.  TEST INFRASTRUCTURE: nothing in the product links, loads or
// runs this program; tests/ and bench.py's CPU leg run it as a checker and a
// timed baseline.
//
// What it restates (same algebra, same order of operations per node):
//   * build (hodlr_routines_omp.h:21-79): bisection mid = (s+e)/2, leaf when
//     len <= leaf, OpenMP tasks for the two children when len > 512, then
//     A12 and A21 of the node compressed (compression.h:67-99);
//   * compression (compression.h:18-99): exact SVD (dgesvd 'S','S') when a
//     block side is below `exact_below` (256 in the reference), otherwise
//     k = rank + oversample columns of Omega ~ U[0,1), Y = A Omega, Q = left
//     singular vectors of Y, B = Q^T A, SVD of B, truncation S_i > tol S_0
//     (rank >= 1, 0 below S_0 <= 1e-15); options the engine exposes on top:
//     a fixed rank and q power iterations (Y = A W, W = orth(A^T orth(Y)));
//   * multiply (hodlr_routines_omp.h:81-135), single or several columns;
//   * solve (hodlr_routines_omp.h:137-240): the recursive Woodbury solve
//     with the capacitance C = [[B12^-1, V_R^T Q2], [V_L^T Q1, B21^-1]],
//     Q_i recomputed by recursive solves on every visit and each leaf
//     refactorised (LU) on every visit, children as OpenMP tasks when the
//     node is longer than 512.
// Deliberate differences: 64-bit indices (the reference's int dimensions
// overflow at N >= 46341, matrix.h:34); the dense A is shared by the tasks
// (the reference's tasks capture `const Matrix&` firstprivate, so libgomp
// copies all of A into every task); B^-1 of the diagonal core is the
// elementwise reciprocal (what dgetrf + dgetri return for a diagonal
// matrix); Omega is drawn per block from mt19937 seeded by (seed, block),
// or read from a file in the serial draw order (hodlr_routines.h:43-53), so
// that results do not depend on task scheduling.
//
// BLAS/LAPACK: the five routines of include/matrix.h:13-26 (dgemm_, dgesvd_,
// dgetrf_, dgetrs_; dgetri_ is not needed, see above) are resolved at run
// time with dlopen from $HMAT_LAPACK (a library exporting them with or
// without a "scipy_" prefix) or MKL's single dynamic library, with one BLAS
// thread as run_tests.sh:7-9 sets.
//
//   hodlr_omp --kernel cauchy|gauss|refcauchy --n N --leaf L --rank r
//             [--oversample p] [--q q] [--fixed 0|1] [--tol t] [--exact-below e]
//             [--sigma s] [--lam l] [--seed s] [--omega file] [--b file]
//             [--out-x file] [--out-y file] [--nrhs c] [--no-solve]
// Prints one JSON line: build / matvec / solve seconds (omp_get_wtime), the
// relative residual ||b - A_H x|| / ||b||, per-depth maximum ranks, threads.
#include <dlfcn.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

// ------------------------------------------------------------ LAPACK ----
using gemm_t = void (*)(const char*, const char*, const int*, const int*, const int*, const double*,
                        const double*, const int*, const double*, const int*, const double*, double*, const int*);
using gesvd_t = void (*)(const char*, const char*, const int*, const int*, double*, const int*, double*, double*,
                         const int*, double*, const int*, double*, const int*, int*);
using getrf_t = void (*)(const int*, const int*, double*, const int*, int*, int*);
using getrs_t = void (*)(const char*, const int*, const int*, const double*, const int*, const int*, double*,
                         const int*, int*);
gemm_t p_gemm;
gesvd_t p_gesvd;
getrf_t p_getrf;
getrs_t p_getrs;
std::string g_lapack;

void* sym(void* h, const char* name) {
  void* f = dlsym(h, name);
  if (!f) f = dlsym(h, (std::string("scipy_") + name).c_str());
  return f;
}

void load_lapack() {
  std::vector<std::string> cands;
  if (const char* e = std::getenv("HMAT_LAPACK")) cands.push_back(e);
  cands.push_back("/opt/conda/lib/libmkl_rt.so.1");
  cands.push_back("libmkl_rt.so.1");
  cands.push_back("liblapack.so.3");
  for (const auto& c : cands) {
    // MKL reads its threading layer when it initialises
    setenv("MKL_THREADING_LAYER", "SEQUENTIAL", 0);
    void* h = dlopen(c.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    p_gemm = (gemm_t)sym(h, "dgemm_");
    p_gesvd = (gesvd_t)sym(h, "dgesvd_");
    p_getrf = (getrf_t)sym(h, "dgetrf_");
    p_getrs = (getrs_t)sym(h, "dgetrs_");
    if (!(p_gemm && p_gesvd && p_getrf && p_getrs)) continue;
    if (auto f = (int (*)(int))dlsym(h, "MKL_Set_Threading_Layer")) f(1);  // sequential
    if (auto f = (void (*)(int))sym(h, "openblas_set_num_threads")) f(1);
    g_lapack = c;
    return;
  }
  throw std::runtime_error("no LAPACK library found (set HMAT_LAPACK)");
}

// -------------------------------------------------------------- dense ----
// Column-major, element (i, j) at v[j * r + i] (matrix.h:36-37), 64-bit sizes.
struct Mat {
  int64_t r = 0, c = 0;
  std::vector<double> v;
  Mat() = default;
  Mat(int64_t rows, int64_t cols) : r(rows), c(cols), v((size_t)(rows * cols), 0.0) {}
  double& operator()(int64_t i, int64_t j) { return v[(size_t)(j * r + i)]; }
  double operator()(int64_t i, int64_t j) const { return v[(size_t)(j * r + i)]; }
  double* p() { return v.data(); }
  const double* p() const { return v.data(); }
};

// C = op(A) op(B)
Mat mul(const Mat& A, bool ta, const Mat& B, bool tb) {
  const int m = (int)(ta ? A.c : A.r), k = (int)(ta ? A.r : A.c), n = (int)(tb ? B.r : B.c);
  Mat C(m, n);
  if (m == 0 || n == 0) return C;
  if (k == 0) return C;
  const double one = 1.0, zero = 0.0;
  const int lda = std::max<int>(1, (int)A.r), ldb = std::max<int>(1, (int)B.r), ldc = std::max(1, m);
  p_gemm(ta ? "T" : "N", tb ? "T" : "N", &m, &n, &k, &one, A.p(), &lda, B.p(), &ldb, &zero, C.p(), &ldc);
  return C;
}

// rows [r0, r0+m) x columns [c0, c0+n) of a column-major array with ld
Mat sub(const double* a, int64_t ld, int64_t r0, int64_t c0, int64_t m, int64_t n) {
  Mat B(m, n);
  for (int64_t j = 0; j < n; ++j) std::memcpy(&B(0, j), a + (c0 + j) * ld + r0, sizeof(double) * m);
  return B;
}

struct Svd {
  Mat U, Vt;
  std::vector<double> S;
};
// dgesvd('S', 'S') (compression.h:25-39)
Svd svd(Mat A) {
  const int m = (int)A.r, n = (int)A.c, mn = std::min(m, n);
  Svd s;
  s.U = Mat(m, mn);
  s.Vt = Mat(mn, n);
  s.S.assign(mn, 0.0);
  if (mn == 0) return s;
  int lda = m, ldu = m, ldvt = mn, info = 0, lwork = -1;
  double q = 0.0;
  p_gesvd("S", "S", &m, &n, A.p(), &lda, s.S.data(), s.U.p(), &ldu, s.Vt.p(), &ldvt, &q, &lwork, &info);
  lwork = (int)q;
  std::vector<double> w(std::max(lwork, 1));
  p_gesvd("S", "S", &m, &n, A.p(), &lda, s.S.data(), s.U.p(), &ldu, s.Vt.p(), &ldvt, w.data(), &lwork, &info);
  if (info != 0) throw std::runtime_error("dgesvd info " + std::to_string(info));
  return s;
}

// X <- A^{-1} X by dgetrf + dgetrs on a copy of A (Matrix::solve_lu,
// matrix.h:120-131, which factorises on every call)
void lu_solve(Mat A, Mat& X) {
  const int n = (int)A.r, nr = (int)X.c;
  if (n == 0 || nr == 0) return;
  std::vector<int> piv(n);
  int info = 0, lda = n, ldb = n;
  p_getrf(&n, &n, A.p(), &lda, piv.data(), &info);
  if (info != 0) throw std::runtime_error("singular block in solve (dgetrf info " + std::to_string(info) + ")");
  p_getrs("N", &n, &nr, A.p(), &lda, piv.data(), X.p(), &ldb, &info);
}

// ------------------------------------------------------- compression ----
struct Opts {
  double tol = 1e-6;
  int rank = 64, oversample = 10, q = 0, exact_below = 256;
  bool fixed = false;
};

bool randomised(int64_t m, int64_t n, const Opts& o) {
  if (m < o.exact_below || n < o.exact_below) return false;
  const int64_t k = std::min<int64_t>(std::min<int64_t>(o.rank + o.oversample, m), n);
  return !(o.fixed && k >= std::min(m, n));
}

int keep_rank(const std::vector<double>& S, const Opts& o) {
  if (o.fixed) return std::min<int>(o.rank, (int)S.size());
  int r = 0;
  if (!S.empty() && S[0] > 1e-15)
    for (double s : S) r += (s > o.tol * S[0]) ? 1 : 0;
  return std::max(r, 1);
}

struct LowRank {
  Mat U, V;               // m x r, n x r
  std::vector<double> S;  // the diagonal core B
};

LowRank truncate(const Svd& s, int r, const Mat* Qleft) {
  LowRank lr;
  const int64_t m = s.U.r, n = s.Vt.c;
  Mat Ur(m, r), Vr(n, r);
  for (int j = 0; j < r; ++j) {
    std::memcpy(&Ur(0, j), s.U.p() + (size_t)j * m, sizeof(double) * m);
    for (int64_t i = 0; i < n; ++i) Vr(i, j) = s.Vt(j, i);
  }
  lr.U = Qleft ? mul(*Qleft, false, Ur, false) : Ur;
  lr.V = Vr;
  lr.S.assign(s.S.begin(), s.S.begin() + r);
  return lr;
}

Mat orth(const Mat& Y) { return svd(Y).U; }  // compress_svd(Y, -1).U: every column

// omega: n x k, U[0,1)
LowRank compress(const Mat& A, const Opts& o, const Mat* omega) {
  if (!randomised(A.r, A.c, o)) {
    Svd s = svd(A);
    return truncate(s, keep_rank(s.S, o), nullptr);
  }
  Mat Y = mul(A, false, *omega, false);
  for (int it = 0; it < o.q; ++it) {
    Mat Z = mul(A, true, orth(Y), false);
    Y = mul(A, false, orth(Z), false);
  }
  Mat Q = orth(Y);
  Svd s = svd(mul(Q, true, A, false));
  return truncate(s, keep_rank(s.S, o), &Q);
}

// --------------------------------------------------------------- tree ----
struct Node {
  int64_t s = 0, e = 0;
  bool leaf = false;
  Node* l = nullptr;
  Node* r = nullptr;
  Mat D, U, V;                    // U, V: this node's basis at its parent
  std::vector<double> S12, S21;   // cores of A12 (left rows) and A21 (right rows)
  int64_t om12 = -1, om21 = -1;   // Omega offsets (serial draw order), -1: exact branch
  int64_t id12 = 0, id21 = 0;     // block indices for per-block Omega streams
};

struct Hodlr {
  Opts o;
  int64_t N = 0, leaf = 0;
  std::deque<Node> pool;  // stable addresses
  Node* root = nullptr;
  const double* A = nullptr;  // dense N x N, ld N
  const double* omega = nullptr;
  uint64_t seed = 1;

  Node* grow(int64_t s, int64_t e) {
    pool.emplace_back();
    Node* n = &pool.back();
    n->s = s;
    n->e = e;
    if (e - s <= leaf) {
      n->leaf = true;
      return n;
    }
    const int64_t mid = (s + e) / 2;
    n->l = grow(s, mid);
    n->r = grow(mid, e);
    return n;
  }
  // the serial class's draw order: post-order, A12 then A21, randomised
  // blocks only (hodlr_routines.h:43-53)
  void number(Node* n, int64_t& pos, int64_t& blk) {
    if (n->leaf) return;
    number(n->l, pos, blk);
    number(n->r, pos, blk);
    const int64_t ml = n->l->e - n->l->s, mr = n->r->e - n->r->s;
    n->id12 = blk++;
    n->id21 = blk++;
    if (randomised(ml, mr, o)) {
      n->om12 = pos;
      pos += mr * std::min<int64_t>(o.rank + o.oversample, std::min(ml, mr));
    }
    if (randomised(mr, ml, o)) {
      n->om21 = pos;
      pos += ml * std::min<int64_t>(o.rank + o.oversample, std::min(ml, mr));
    }
  }
  Mat omega_block(int64_t off, int64_t id, int64_t n, int64_t k) const {
    Mat W(n, k);
    if (omega) {
      std::memcpy(W.p(), omega + off, sizeof(double) * n * k);
    } else {
      std::mt19937 g((uint32_t)(seed * 1000003u + (uint64_t)id));
      std::uniform_real_distribution<double> u(0.0, 1.0);
      for (auto& v : W.v) v = u(g);
    }
    return W;
  }

  void build_node(Node* n) {
    const int64_t len = n->e - n->s;
    if (n->leaf) {
      n->D = sub(A, N, n->s, n->s, len, len);
      return;
    }
    if (len > 512) {
#pragma omp task
      build_node(n->l);
#pragma omp task
      build_node(n->r);
    } else {
      build_node(n->l);
      build_node(n->r);
    }
    const int64_t mid = n->l->e, ml = mid - n->s, mr = n->e - mid;
    const int64_t k = std::min<int64_t>(o.rank + o.oversample, std::min(ml, mr));
    Mat A12 = sub(A, N, n->s, mid, ml, mr), A21 = sub(A, N, mid, n->s, mr, ml);
    Mat W12, W21;
    if (n->om12 >= 0) W12 = omega_block(n->om12, n->id12, mr, k);
    if (n->om21 >= 0) W21 = omega_block(n->om21, n->id21, ml, k);
    LowRank a = compress(A12, o, n->om12 >= 0 ? &W12 : nullptr);
    LowRank b = compress(A21, o, n->om21 >= 0 ? &W21 : nullptr);
    n->l->U = std::move(a.U);
    n->r->V = std::move(a.V);
    n->S12 = std::move(a.S);
    n->r->U = std::move(b.U);
    n->l->V = std::move(b.V);
    n->S21 = std::move(b.S);
    if (len > 512) {
#pragma omp taskwait
    }
  }

  int64_t omega_length() {
    int64_t pos = 0, blk = 0;
    number(root, pos, blk);
    return pos;
  }

  // om: Omega of every randomised block in the serial draw order (om_len
  // doubles), or null for per-block generated Omega
  void build(const double* a, int64_t n, int64_t lf, const double* om, int64_t om_len) {
    A = a;
    N = n;
    leaf = lf;
    omega = om;
    pool.clear();
    root = grow(0, n);
    const int64_t need = omega_length();
    if (om && om_len != need)
      throw std::runtime_error("omega holds " + std::to_string(om_len) + " doubles, the schedule needs " +
                               std::to_string(need));
#pragma omp parallel
#pragma omp single
    build_node(root);
  }

  // y (len x c) = A_node x (len x c)
  Mat mult(const Node* n, const Mat& x) const {
    const int64_t len = n->e - n->s, c = x.c;
    if (n->leaf) return mul(n->D, false, x, false);
    const int64_t ml = n->l->e - n->s, mr = len - ml;
    Mat xl = sub(x.p(), x.r, 0, 0, ml, c), xr = sub(x.p(), x.r, ml, 0, mr, c);
    Mat yl, yr;
    if (len > 512) {
#pragma omp task shared(yl, xl)
      yl = mult(n->l, xl);
#pragma omp task shared(yr, xr)
      yr = mult(n->r, xr);
#pragma omp taskwait
    } else {
      yl = mult(n->l, xl);
      yr = mult(n->r, xr);
    }
    Mat t1 = mul(n->r->V, true, xr, false), t3 = mul(n->l->V, true, xl, false);
    for (int64_t j = 0; j < c; ++j) {
      for (size_t i = 0; i < n->S12.size(); ++i) t1(i, j) *= n->S12[i];
      for (size_t i = 0; i < n->S21.size(); ++i) t3(i, j) *= n->S21[i];
    }
    Mat tl = mul(n->l->U, false, t1, false), tr = mul(n->r->U, false, t3, false);
    Mat y(len, c);
    for (int64_t j = 0; j < c; ++j) {
      for (int64_t i = 0; i < ml; ++i) y(i, j) = yl(i, j) + tl(i, j);
      for (int64_t i = 0; i < mr; ++i) y(ml + i, j) = yr(i, j) + tr(i, j);
    }
    return y;
  }

  // x = A_node^{-1} b: the recursive Woodbury solve (hodlr_routines_omp.h:147-240)
  Mat solve(const Node* n, const Mat& b) const {
    const int64_t len = n->e - n->s, c = b.c;
    if (n->leaf) {
      Mat x = b;
      lu_solve(n->D, x);
      return x;
    }
    const int64_t ml = n->l->e - n->s, mr = len - ml;
    Mat b1 = sub(b.p(), b.r, 0, 0, ml, c), b2 = sub(b.p(), b.r, ml, 0, mr, c);
    Mat z1, z2, Q1, Q2;
    const bool spawn = len > 512;
    if (spawn) {
#pragma omp task shared(z1, b1)
      z1 = solve(n->l, b1);
#pragma omp task shared(z2, b2)
      z2 = solve(n->r, b2);
#pragma omp taskwait
#pragma omp task shared(Q1)
      Q1 = solve(n->l, n->l->U);
#pragma omp task shared(Q2)
      Q2 = solve(n->r, n->r->U);
#pragma omp taskwait
    } else {
      z1 = solve(n->l, b1);
      z2 = solve(n->r, b2);
      Q1 = solve(n->l, n->l->U);
      Q2 = solve(n->r, n->r->U);
    }
    const int64_t r1 = (int64_t)n->S12.size(), r2 = (int64_t)n->S21.size(), R = r1 + r2;
    Mat C(R, R);
    for (int64_t i = 0; i < r1; ++i) C(i, i) = 1.0 / n->S12[i];
    for (int64_t i = 0; i < r2; ++i) C(r1 + i, r1 + i) = 1.0 / n->S21[i];
    Mat TR = mul(n->r->V, true, Q2, false), BL = mul(n->l->V, true, Q1, false);
    for (int64_t j = 0; j < r2; ++j)
      for (int64_t i = 0; i < r1; ++i) C(i, r1 + j) = TR(i, j);
    for (int64_t j = 0; j < r1; ++j)
      for (int64_t i = 0; i < r2; ++i) C(r1 + i, j) = BL(i, j);
    Mat vz1 = mul(n->r->V, true, z2, false), vz2 = mul(n->l->V, true, z1, false);
    Mat g(R, c);
    for (int64_t j = 0; j < c; ++j) {
      for (int64_t i = 0; i < r1; ++i) g(i, j) = vz1(i, j);
      for (int64_t i = 0; i < r2; ++i) g(r1 + i, j) = vz2(i, j);
    }
    lu_solve(C, g);
    Mat g1 = sub(g.p(), R, 0, 0, r1, c), g2 = sub(g.p(), R, r1, 0, r2, c);
    Mat c1 = mul(Q1, false, g1, false), c2 = mul(Q2, false, g2, false);
    Mat x(len, c);
    for (int64_t j = 0; j < c; ++j) {
      for (int64_t i = 0; i < ml; ++i) x(i, j) = z1(i, j) - c1(i, j);
      for (int64_t i = 0; i < mr; ++i) x(ml + i, j) = z2(i, j) - c2(i, j);
    }
    return x;
  }

  Mat multiply(const Mat& x) const {
    Mat y;
#pragma omp parallel
#pragma omp single
    y = mult(root, x);
    return y;
  }
  Mat solve(const Mat& b) const {
    Mat x;
#pragma omp parallel
#pragma omp single
    x = solve(root, b);
    return x;
  }
  void max_ranks(const Node* n, int d, std::vector<int>& out) const {
    if (n->leaf) return;
    if ((int)out.size() <= d) out.resize(d + 1, 0);
    out[d] = std::max<int>(out[d], (int)std::max(n->S12.size(), n->S21.size()));
    max_ranks(n->l, d + 1, out);
    max_ranks(n->r, d + 1, out);
  }
};

std::vector<double> read_file(const std::string& path, size_t n) {
  std::vector<double> v(n);
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  const size_t got = std::fread(v.data(), sizeof(double), n, f);
  std::fclose(f);
  if (got != n) throw std::runtime_error(path + ": expected " + std::to_string(n) + " doubles");
  return v;
}
void write_file(const std::string& path, const Mat& m) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fwrite(m.p(), sizeof(double), m.v.size(), f);
  std::fclose(f);
}
int64_t file_doubles(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) throw std::runtime_error("cannot open " + path);
  std::fseek(f, 0, SEEK_END);
  const long b = std::ftell(f);
  std::fclose(f);
  return b / (long)sizeof(double);
}

}  // namespace

int main(int argc, char** argv) {
  try {
    std::string kernel = "cauchy", omega_path, b_path, outx, outy;
    int64_t N = 4096, leaf = 256, nrhs = 1;
    double sigma = 0.1, lam = 1.0;
    uint64_t seed = 1;
    bool do_solve = true;
    Opts o;
    for (int i = 1; i < argc; ++i) {
      const std::string a = argv[i];
      auto val = [&]() -> std::string {
        if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
        return argv[++i];
      };
      if (a == "--kernel") kernel = val();
      else if (a == "--n") N = std::stoll(val());
      else if (a == "--leaf") leaf = std::stoll(val());
      else if (a == "--rank") o.rank = std::stoi(val());
      else if (a == "--oversample") o.oversample = std::stoi(val());
      else if (a == "--q") o.q = std::stoi(val());
      else if (a == "--fixed") o.fixed = std::stoi(val()) != 0;
      else if (a == "--tol") o.tol = std::stod(val());
      else if (a == "--exact-below") o.exact_below = std::stoi(val());
      else if (a == "--sigma") sigma = std::stod(val());
      else if (a == "--lam") lam = std::stod(val());
      else if (a == "--seed") seed = std::stoull(val());
      else if (a == "--omega") omega_path = val();
      else if (a == "--b") b_path = val();
      else if (a == "--nrhs") nrhs = std::stoll(val());
      else if (a == "--out-x") outx = val();
      else if (a == "--out-y") outy = val();
      else if (a == "--no-solve") do_solve = false;
      else throw std::runtime_error("unknown argument " + a);
    }
    load_lapack();
    {  // first calls initialise the library (dispatch tables, buffers): not timed
      Mat w(64, 64);
      for (int64_t i = 0; i < 64; ++i) w(i, i) = 1.0 + (double)i;
      (void)svd(mul(w, false, w, true));
    }
    // dense A (untimed, as in the reference drivers, src/main_hodlr.cpp:48-49)
    std::vector<double> A((size_t)(N * N));
    std::vector<double> pts(N);
    if (kernel == "gauss") {
      std::mt19937_64 g(seed);
      std::uniform_real_distribution<double> u(0.0, 1.0);
      for (auto& p : pts) p = u(g);
      std::sort(pts.begin(), pts.end());
    }
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < N; ++j)
      for (int64_t i = 0; i < N; ++i) {
        double v;
        if (kernel == "cauchy") {
          v = 1.0 / ((double)i / N - ((double)j + 0.5) / N);  // 1/(x_i - y_j)
        } else if (kernel == "gauss") {
          const double d = pts[i] - pts[j];
          v = std::exp(-d * d / (2.0 * sigma * sigma)) + (i == j ? lam : 0.0);
        } else {  // the drivers' Cauchy matrix, kernel.h:30-48 with shift 1
          v = 1.0 / ((double)i / (N - 1) + (double)j / (N - 1) + 1.0) + (i == j ? 1.0 : 0.0);
        }
        A[(size_t)(j * N + i)] = v;
      }
    Hodlr h;
    h.o = o;
    h.seed = seed;
    std::vector<double> om;
    if (!omega_path.empty()) om = read_file(omega_path, (size_t)file_doubles(omega_path));
    Mat b(N, nrhs);
    if (!b_path.empty()) {
      b.v = read_file(b_path, (size_t)(N * nrhs));
    } else {
      std::mt19937_64 g(seed + 7);
      std::normal_distribution<double> nd(0.0, 1.0);
      for (auto& v : b.v) v = nd(g);
    }
    double t0 = omp_get_wtime();
    h.build(A.data(), N, leaf, om.empty() ? nullptr : om.data(), (int64_t)om.size());
    const double t_build = omp_get_wtime() - t0;
    const int64_t olen = h.omega_length();
    t0 = omp_get_wtime();
    Mat y = h.multiply(b);
    const double t_mv = omp_get_wtime() - t0;
    double t_solve = 0.0, resid = -1.0;
    Mat x;
    if (do_solve) {
      t0 = omp_get_wtime();
      x = h.solve(b);
      t_solve = omp_get_wtime() - t0;
      Mat ax = h.multiply(x);
      double num = 0.0, den = 0.0;
      for (size_t i = 0; i < b.v.size(); ++i) {
        num += (b.v[i] - ax.v[i]) * (b.v[i] - ax.v[i]);
        den += b.v[i] * b.v[i];
      }
      resid = std::sqrt(num / den);
    }
    if (!outx.empty() && do_solve) write_file(outx, x);
    if (!outy.empty()) write_file(outy, y);
    std::vector<int> ranks;
    h.max_ranks(h.root, 0, ranks);
    std::string rk;
    for (size_t i = 0; i < ranks.size(); ++i) rk += (i ? "," : "") + std::to_string(ranks[i]);
    std::printf(
        "{\"n\": %lld, \"leaf\": %lld, \"build_s\": %.6f, \"matvec_s\": %.6f, \"solve_s\": %.6f, \"resid\": %.3e, "
        "\"threads\": %d, \"max_rank_per_depth\": [%s], \"omega_length\": %lld, \"lapack\": \"%s\"}\n",
        (long long)N, (long long)leaf, t_build, t_mv, t_solve, resid, omp_get_max_threads(), rk.c_str(),
        (long long)olen, g_lapack.c_str());
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "hodlr_omp: %s\n", e.what());
    return 1;
  }
}
