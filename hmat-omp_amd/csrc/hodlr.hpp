// HODLR engine: the MI355X counterpart of HODLRMatrixOMP
// (include/hodlr_routines_omp.h:9-241).
#pragma once
#include <vector>

#include "engine.hpp"

namespace hmat {

enum class BlockMode { RSVD, ExactWide, ExactTall };

// Does an m x n block take the randomised branch?  compression.h:72-75 in
// tolerance mode; in fixed-rank mode also exact when k >= min(m, n).
inline bool uses_rsvd(int64_t m, int64_t n, const hmat_options& o) {
  if (m < o.exact_below || n < o.exact_below) return false;
  const int64_t k = std::min<int64_t>(std::min<int64_t>(o.rank + o.oversample, m), n);
  if (o.fixed_rank && k >= std::min(m, n)) return false;
  return true;
}
inline int sketch_width(int64_t m, int64_t n, const hmat_options& o) {
  return (int)std::min<int64_t>(std::min<int64_t>(o.rank + o.oversample, m), n);
}

// Off-diagonal blocks of one tree depth c >= 1.  Segment s (a node at depth
// c) is the row set of the block A(s, sib(s)).
struct HLevel {
  std::vector<int> ids;    // node ids (start order)
  std::vector<int> sib;    // position of the sibling
  std::vector<int> rank;   // rank of block (s, sib(s))
  int rmax = 0;
  DevBuf<double> US;       // N x rmax: U_s diag(S_s) at the rows of s
  DevBuf<double> V;        // N x rmax: V of block (sib(s), s) at the rows of s
  std::vector<double> S;   // nseg x rmax singular values (host copy)
};

// Internal nodes of depth j with their Woodbury capacitance inverses.
struct HParents {
  std::vector<int> ids, lpos, rpos;
  std::vector<int> msize;
  int mmax = 0;
  DevBuf<double> Minv;  // np x mmax x mmax
};

class HodlrEngine {
 public:
  explicit HodlrEngine(const hmat_options& o);
  ~HodlrEngine();

  void build(const GenDesc& g, int64_t N, int leaf, const double* host_omega, int64_t omega_len);
  void matvec(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy);
  void factor();
  void solve(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);
  void to_dense_host(double* A, int64_t lda);

  static std::vector<std::pair<int64_t, int>> omega_schedule(const hmat_options& o, int64_t N,
                                                             int leaf);

  hmat_options opt;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int64_t N = 0;
  int leaf = 0;
  Tree tree;
  std::vector<HLevel> lv;  // index = depth (lv[0] unused)
  DevBuf<double> D, Dinv;  // nleaf x lmax x lmax
  bool built = false, factored = false;
  std::vector<HParents> par;  // index = depth of the parent
  DevBuf<double> Wall;        // N x Rtot: Q_c = A_node^{-1} U_c S_c
  std::vector<int64_t> colofs;
  int64_t Rtot = 0;
  double refine_rtol = 1e-10;
  int refine_max = 50;
  int last_iters = 0;
  double last_resid = 0.0;

 private:
  void build_level(int c, const GenDesc& g, const std::vector<int64_t>& omega_off,
                   const double* host_omega);
  void correction(int j, double* W, int64_t ldw, int ncols);
  void leaf_apply(const DevBuf<double>& Dm, const double* B, int64_t ldb, int ncols, double* Y,
                  int64_t ldy, double beta);

  kern::Workspace ws_;
  Orthogonalizer orth_;
  DevBuf<SegNN> dnn_;
  DevBuf<SegTN> dtn_;
  DevBuf<SegSk> dsk_;
  DevBuf<SmallMat> dsm_;
  DevBuf<BlockDesc> dbd_;
  DevBuf<double> Yb_, Xb_, Tb_, Rb_, Vr_, Sv_, Wk_, TU_, TV_;
  DevBuf<double> tvec_, coef_, wtmp_;
  DevBuf<int> info_;
};

}  // namespace hmat
