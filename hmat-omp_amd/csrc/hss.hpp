// HSS engine: the MI355X counterpart of HSSMatrixOMP
// (include/hss_routines_omp.h:11-375).
#pragma once
#include <vector>

#include "hodlr.hpp"

namespace hmat {

class HssEngine {
 public:
  explicit HssEngine(const hmat_options& o);
  ~HssEngine();

  void build(const GenDesc& g, int64_t N, int leaf, bool nested, const double* host_omega,
             int64_t omega_len);
  void matvec(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy);
  void factor();
  void solve(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);
  void to_dense_host(double* A, int64_t lda);
  int root_rank() const;
  static int64_t omega_length(const hmat_options& o, int64_t N, int leaf);

  hmat_options opt;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int64_t N = 0;
  int leaf = 0;
  bool built = false, factored = false;
  DevBuf<double> px, py;  // kernel points
  Tree tree;
};

}  // namespace hmat
