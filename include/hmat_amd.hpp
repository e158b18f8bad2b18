// hmat_amd.hpp — C++ drop-in for the OpenMP classes of chenxinye/hmat-omp,
// running on the MI355X engine behind the C ABI of hmat_amd.h.
//
//   HODLRMatrixOMPT<M>  replaces HODLRMatrixOMP (include/hodlr_routines_omp.h:9-241)
//   HSSMatrixOMPT<M>    replaces HSSMatrixOMP   (include/hss_routines_omp.h:11-375)
//   HODLRMatrixT<M>, HSSMatrixT<M>  replace the serial HODLRMatrix / HSSMatrix
// (include/hmat_amd_dropin.hpp adds the reference's Matrix, Kernels:: and
// the four class names, for drivers written against the reference)
//
// M is the caller's dense matrix type.  It needs what the reference's Matrix
// (include/matrix.h:28-150) provides: M(rows, cols) zero-initialised,
// rows(), cols(), and data() with column-major storage (element (i, j) at
// data()[j * rows() + i]).  The reference's own Matrix satisfies this, so a
// driver written against the reference keeps compiling with
//
//   #include "matrix.h"      // the caller's Matrix
//   #include "hmat_amd.hpp"
//   using HODLRMatrixOMP = hmat_amd::HODLRMatrixOMPT<Matrix>;
//   using HSSMatrixOMP = hmat_amd::HSSMatrixOMPT<Matrix>;
//
// Same members and argument meaning as the reference:
//   constructor(tol) with the same defaults (1e-6 HODLR, 1e-9 HSS), public
//   N / tolerance / rsvd_rank, build_from_dense(A, leaf_size), multiply(x)
//   and solve(b).  As in the reference, multiply uses column 0 of x and
//   returns N x 1 (hodlr_routines_omp.h:81-90, hss_routines_omp.h:154-160);
//   HODLR solve takes every column of b (hodlr_routines_omp.h:137-145) and
//   HSS solve column 0 (hss_routines_omp.h:235-245).
// Differences: the tree is on the device, so the public root / nodes
// pointers of the reference do not exist; errors that the reference leaves
// to assert or silently ignores (LAPACK info codes) throw hmat_amd::Error
// carrying the hmat_status.  The Omega of every randomised block is drawn on
// the device from options.seed (the reference draws from std::random_device).
#ifndef HMAT_AMD_HPP
#define HMAT_AMD_HPP

#include <stdexcept>
#include <string>

#include "hmat_amd.h"

namespace hmat_amd {

class Error : public std::runtime_error {
 public:
  Error(hmat_status s, const std::string& what) : std::runtime_error(what), status(s) {}
  hmat_status status;
};

inline void check(hmat_status s) {
  if (s != HMAT_OK) throw Error(s, hmat_last_error());
}

// Shared part of the two classes: options, handle lifetime, host copies.
template <class H, class Ops>
class HandleBase {
 public:
  int N = 0;
  double tolerance;
  int rsvd_rank = 64;
  hmat_options options;  // set before build_from_dense to change the rule

  explicit HandleBase(double tol) : tolerance(tol) {
    hmat_options_reference(&options, tol);
    check(Ops::create(&h_, &options));
  }
  ~HandleBase() {
    if (h_) Ops::destroy(h_);
  }
  HandleBase(const HandleBase&) = delete;
  HandleBase& operator=(const HandleBase&) = delete;

 protected:
  // tolerance / rsvd_rank are public fields in the reference and may be
  // changed between construction and build: rebuild the handle if so
  void sync() {
    if (options.tol != tolerance || options.rank != rsvd_rank) {
      options.tol = tolerance;
      options.rank = rsvd_rank;
      if (h_) Ops::destroy(h_);
      h_ = nullptr;
      check(Ops::create(&h_, &options));
    }
  }
  H h_ = nullptr;
};

struct HodlrOps {
  static hmat_status create(hmat_hodlr* h, const hmat_options* o) { return hmat_hodlr_create(h, o); }
  static hmat_status destroy(hmat_hodlr h) { return hmat_hodlr_destroy(h); }
};
struct HssOps {
  static hmat_status create(hmat_hss* h, const hmat_options* o) { return hmat_hss_create(h, o); }
  static hmat_status destroy(hmat_hss h) { return hmat_hss_destroy(h); }
};

template <class M>
class HODLRMatrixOMPT : public HandleBase<hmat_hodlr, HodlrOps> {
 public:
  explicit HODLRMatrixOMPT(double tol = 1e-6) : HandleBase(tol) {}

  // hodlr_routines_omp.h:21-31
  void build_from_dense(const M& A, int leaf_size) {
    sync();
    N = A.rows();
    check(hmat_hodlr_build_dense(h_, A.data(), A.rows(), A.rows(), leaf_size, nullptr, 0));
  }
  // hodlr_routines_omp.h:81-90 (column 0 of x)
  M multiply(const M& x) {
    M y(N, 1);
    check(hmat_hodlr_matvec(h_, x.data(), x.rows(), 1, y.data(), N));
    return y;
  }
  // hodlr_routines_omp.h:137-145 (every column of b)
  M solve(const M& b) {
    M x(N, b.cols());
    check(hmat_hodlr_solve(h_, b.data(), b.rows(), b.cols(), x.data(), N));
    return x;
  }
  hmat_hodlr handle() const { return h_; }
};

template <class M>
class HSSMatrixOMPT : public HandleBase<hmat_hss, HssOps> {
 public:
  // mode HMAT_HSS_REFERENCE (0, false, the default) reproduces the reference
  // operator (identity transfers, hss_routines_omp.h:110-119);
  // HMAT_HSS_NESTED (1, true) compresses the transfers; HMAT_HSS_SAMPLED (2)
  // builds the nested form from one sample pass per side, O(N s) memory.
  explicit HSSMatrixOMPT(double tol = 1e-9, int mode = HMAT_HSS_REFERENCE) : HandleBase(tol), nested_(mode) {}

  // hss_routines_omp.h:62-149
  void build_from_dense(const M& A, int leaf_size) {
    sync();
    N = A.rows();
    check(hmat_hss_build_dense(h_, A.data(), A.rows(), A.rows(), leaf_size, nested_, nullptr, 0));
  }
  // hss_routines_omp.h:154-160 (column 0 of x)
  M multiply(const M& x) {
    M y(N, 1);
    check(hmat_hss_matvec(h_, x.data(), x.rows(), 1, y.data(), N));
    return y;
  }
  // hss_routines_omp.h:235-245 (column 0 of b)
  M solve(const M& b) {
    M x(N, 1);
    check(hmat_hss_solve(h_, b.data(), b.rows(), 1, x.data(), N));
    return x;
  }
  hmat_hss handle() const { return h_; }

 private:
  int nested_;
};

// The serial classes of the reference compute the same algebra as the
// OpenMP ones (the engine is one GPU either way):
//   HODLRMatrix  include/hodlr_routines.h:8-176 (solve takes every column)
//   HSSMatrix    include/hss_routines.h:10-307  (solve takes every column,
//                solve_multi_rhs, hss_routines.h:197-199 and 281-290)
template <class M>
using HODLRMatrixT = HODLRMatrixOMPT<M>;

template <class M>
class HSSMatrixT : public HSSMatrixOMPT<M> {
 public:
  explicit HSSMatrixT(double tol = 1e-9, int mode = HMAT_HSS_REFERENCE) : HSSMatrixOMPT<M>(tol, mode) {}
  // hss_routines.h:197-199 (every column of b)
  M solve(const M& b) {
    M x(this->N, b.cols());
    check(hmat_hss_solve(this->h_, b.data(), b.rows(), b.cols(), x.data(), this->N));
    return x;
  }
};

}  // namespace hmat_amd

#endif  // HMAT_AMD_HPP
