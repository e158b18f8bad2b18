// HSS engine: the MI355X counterpart of HSSMatrixOMP
// (include/hss_routines_omp.h:11-375).
//
// Generators (per node p with children a, b):
//   leaves:   D_i, U_i (len x rU_i), V_i (len x rV_i) — the reference's
//             leaf bases: RSVD of the whole off-diagonal row block A(I, I^c)
//             and column block A(I^c, I)^T (hss_routines_omp.h:69-95);
//   internal: transfers R_p ((rU_a + rU_b) x rU_p), W_p (same for V), and
//             couplings B_ab = U_a^T A(I_a, I_b) V_b, B_ba = U_b^T A(I_b, I_a) V_a
//             with the full (nested) bases of the children.
// Build modes: kReference reproduces the reference operator: R = W = I and
// ranks add up (hss_routines_omp.h:110-119).  kNested compresses the
// transfers (tolerance or fixed rank), so ranks stay bounded and the
// operator costs O(N r) to apply; its leaf bases still come from the N-wide
// row / column blocks, O(N^2 k / leaf) build memory.  kSampled builds the
// same nested form from one sample pass per side, Y = A Omega and
// Z = A^T Psi (Martinsson 2011, orthonormal bases): leaf bases from the
// samples with the leaf's own block excluded, parents from the children's
// reduced samples, couplings from one generated-entry product per depth;
// O(N k) build memory.
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "hodlr.hpp"
#include "hss_mv.hpp"

namespace hmat {

struct SNode {
  int rU = 0, rV = 0;
  // offsets into the small-matrix pool and their leading dimensions
  // (uniform over a depth so that one launch covers the depth)
  int64_t oR = -1, oW = -1;      // transfers: (rU_a+rU_b) x rU, (rV_a+rV_b) x rV
  int64_t oBab = -1, oBba = -1;  // couplings: rU_a x rV_b, rU_b x rV_a
  int ldR = 1, ldW = 1, ldB = 1;
  int64_t oMinv = -1, oG = -1, oK = -1;  // factor: M^{-1} (ms x ms), G (kR x rU), K_hat (rV x rU)
  int64_t oM = -1;                        // factor: M itself, for the refinement step of M solves
  int ms = 0;                             // rV_a + rV_b
};

class HssEngine {
 public:
  explicit HssEngine(const hmat_options& o);
  ~HssEngine();

  enum Mode { kReference = 0, kNested = 1, kSampled = 2 };
  // factor()/solve(): the reference's push-through Woodbury recursion, or a
  // ULV factorisation (orthogonal eliminations, backward stable; nested
  // bases with every reduced block at most 1024 rows)
  enum Solver { kWoodbury = 0, kUlv = 1 };
  void build(const GenDesc& g, int64_t N, int leaf, int mode, const double* host_omega,
             int64_t omega_len);
  void matvec(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy);
  void factor();
  void solve(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);
  void to_dense_host(double* A, int64_t lda);
  // operator file: options, tree, per-node ranks and generator offsets,
  // transfers/couplings, leaf bases and leaf blocks (no factorisation)
  void save(const std::string& path);
  void load(const std::string& path);
  int root_rank() const;
  // HBM bytes one matvec / one apply-inverse must move at least: leaf
  // blocks (and, for the refined leaf solve, their inverses), the leaf bases
  // at their ranks, the small generator (and factor) matrices, and the
  // right-hand sides in and solutions out
  double matvec_bytes(int nrhs) const;
  double solve_bytes(int nrhs) const;
  static int64_t omega_length(const hmat_options& o, int64_t N, int leaf);
  // sample width of the sampled build (rank + oversample, at most N) and
  // its host Omega length: [Omega | Psi], each N x width, column-major
  static int sampled_width(const hmat_options& o, int64_t N);
  static int64_t sampled_omega_length(const hmat_options& o, int64_t N);

  hmat_options opt;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int64_t N = 0;
  int leaf = 0;
  bool nested = false;
  int mode = kReference;
  bool built = false, factored = false;
  int solver = kWoodbury;
  // FP64 iterative refinement of solves through the HSS matvec (off by
  // default: rtol 0); iterations and max relative residual of the last solve
  double refine_rtol = 0.0;
  int refine_max = 30;
  int last_iters = 0;
  double last_resid = -1.0;
  DevBuf<double> px, py;  // kernel points
  Tree tree;
  std::vector<SNode> sn;  // per tree node
  int rUmax = 0, rVmax = 0;

 private:
  void compress_leaves(const GenDesc& g, bool col_side, const double* host_omega,
                       const std::vector<int64_t>& omega_off);
  void build_internal();
  void build_sampled(const GenDesc& g, const double* host_omega, int64_t omega_len);
  int64_t spool_alloc(int64_t n);
  int64_t pool_alloc(int64_t n);
  void set_ldv();
  void leaf_apply(const double* Dm, const double* B, int64_t ldb, int ncols, double* Y, int64_t ldy,
                  double alpha = 1.0, double beta = 0.0);
  // Y = D^{-1} B over all leaves: the inverse, then one refinement step
  // against D (residual in R, N x ncols, ld N), as the HODLR factorisation
  // does: backward stable like an LU solve while every step stays a GEMM
  void leaf_solve(const double* B, int64_t ldb, int ncols, double* Y, int64_t ldy, double* R);
  void apply_inverse(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);
  void plan_apply(int nrhs);
  // descriptors of the one-launch matvec; false when the operator's shape
  // is beyond it (ranks above 64, leaves above 1024 rows, a single leaf) or
  // the device cannot hold its grid
  bool plan_matvec();
  void ulv_factor();
  void ulv_apply(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);
  double ulv_bytes(int nrhs) const;

  // leaf generators
  DevBuf<double> D, Dinv, Ul, Vl, Ql;  // D/Dinv: nleaf x lmax^2; Ul, Vl, Ql: N x rmax
  // projected blocks of the leaves (row side T^T, column side S^T), one
  // N x r column block per leaf, and each leaf's column offset
  DevBuf<double> leafT[2];
  std::vector<int64_t> leafcol[2];
  // small matrices (transfers, couplings, factors) in one device pool
  DevBuf<double> pool;
  int64_t pool_used = 0;
  int64_t pool_built = 0;  // end of the build's part (the factor's follows)
  int ldv = 1;  // leading dimension of every per-node vector in the sweeps
  DevBuf<double> va_, vb_, vc_, ve_, vf_;
  DevBuf<double> lres_;  // leaf-solve residual
  DevBuf<double> ir_y_, ir_r_, ir_n_, ir_p_;  // refinement: A x, residual, norms, norm partials
  DevBuf<double> ir_best_;                     // the iterate before the last correction
  // Build scratch, kept for the next build and grown only when a build
  // needs more (freeing and reallocating multi-GB arrays per build stalls
  // the host for up to a second per array).
  struct LeafScratch {  // compress_leaves
    DevBuf<double> Yb, Tb, Bw, Tw2, X, Rb, Vr, Sv, Wk, TU, TV;
  } ls_;
  struct NodeScratch {  // build_internal: wide arrays of the child and the current depth
    DevBuf<double> T[2], S[2], VF[2], Pt, Ps, Tmp, Pn[2], Rb, Vr, Sv, Wk, TU, TV;
  } ns_;
  struct SampleScratch {  // build_sampled
    DevBuf<double> Om, Ps, Y, Z, Tb, Rt, Sm, Vr, Sv, Wk, TU, TV, P, Mt;
    DevBuf<double> UF[2], VF[2];  // full bases of the child depth and the current one
    DevBuf<double> sp;            // reduced samples per node (see build_sampled)
    int64_t sp_used = 0;
  } ss_;

  // ULV factor, per node: n block rows (leaf size, or the children's ranks
  // summed), r = rank of U, m = n - r rows eliminated, s = rank of V; offsets
  // into upool_ of the Householder vectors of U (n x r) with their tau and
  // signs, of C[r:]^T = H_P [Rq; 0] (n x m, reflectors below Rq) with tau,
  // of G = H_P^T C[:r]^T (n x r), of H_P^T V (n x s) and of Dt (r x r); all
  // leading dimensions n (Dt: r)
  struct UNode {
    int n = 0, r = 0, m = 0, s = 0;
    int64_t oHU = -1, oTU = -1, oSU = -1, oHP = -1, oTP = -1, oG = -1, oV = -1, oDt = -1;
  };
  std::vector<UNode> un_;
  DevBuf<double> upool_, uC_;           // factor; reduced blocks of one depth
  DevBuf<double> ubeta_, ugk_, uzero_;  // solve: right-hand sides and known parts per node
  DevBuf<int> uinfo_;

  kern::HmvPlan mv_;  // one-launch matvec (valid for the built operator)
  // descriptors of the Woodbury apply-inverse for one column count (valid
  // for the factorisation and the per-node vectors they were built for)
  struct ApplyPlan {
    int nrhs = 0;
    bool ok = false;
    kern::NNPlan leaf, expand;
    kern::TNPlan vt;
    DevBuf<kern::BGemm> ops;
    DevBuf<kern::NodeTask> tasks;
    std::vector<std::pair<int, int>> phases;
    std::vector<const void*> addrs;
  } apl_;
  // the ULV apply-inverse's launches for one column count and one pair of
  // right-hand side / solution buffers: descriptors on the device, replayed
  // by a list of steps (0: reflector applications, 1: triangular solves,
  // 2: node tasks; rows and cols size the launch)
  struct UlvStep {
    int32_t kind, off, count, rows, cols;
  };
  struct UlvPlan {
    bool ok = false;
    int nrhs = 0;
    std::vector<const void*> key;
    std::vector<UlvStep> steps;
    DevBuf<unsigned char> items;  // reflector descriptors, then (at tr_off) triangular-solve descriptors
    int64_t tr_off = 0;
    DevBuf<kern::BGemm> ops;
    DevBuf<kern::NodeTask> tasks;
  } ulvp_;
  kern::Workspace ws_;
  Orthogonalizer orth_;
  DevBuf<SegNN> dnn_;
  DevBuf<SmallMat> dsm_;
  DevBuf<BlockDesc> dbd_;
  DevBuf<kern::BGemm> dbg_;
  DevBuf<kern::ZeroRect> dzr_;
  DevBuf<kern::Rect> drc_;
  DevBuf<int> info_;
};

}  // namespace hmat
