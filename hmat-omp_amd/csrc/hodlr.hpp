// HODLR engine: the MI355X counterpart of HODLRMatrixOMP
// (include/hodlr_routines_omp.h:9-241).
#pragma once
#include <map>
#include <memory>
#include <type_traits>
#include <string>
#include <vector>

#include "apply.hpp"
#include "engine.hpp"
#include "matvec.hpp"

namespace hmat {

enum class BlockMode { RSVD, ExactWide, ExactTall };

// BF16 terms per operand in the mixed-precision range finder.  The pass
// whose output is orthonormalised into the final basis Q uses 3 terms (the
// 24 bits of an FP32 entry); the power-iteration passes before it use one.
// Subspace iteration damps an earlier pass's error E by the gap
// sigma_{k+1}/sigma_r before it reaches Q, so the accuracy of Q is set by the
// last pass: for a 4096^2 Cauchy block, k = 58, r = 48, q = 2, the
// approximation error is 3.2e-8 with BF16 power passes and 2.9e-8 with all
// passes BF16x3 (1.7e-4 with all passes single BF16).
constexpr int kLowPrecSplit = 3;
constexpr int kPowerSplit = 1;
// Jacobi rotation floor of the mixed path, (2^-23)^2 of ||R||_F^2.
constexpr double kMixedJacobiFloor = 1.4210854715202004e-14;
// Scratch budget of one group of depths built together (three N x k
// arrays per depth).
constexpr double kGroupScratchBytes = 8e9;

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

// Internal nodes of depth j (children at depth j+1).
struct HParents {
  std::vector<int> ids, lpos, rpos;
  std::vector<int> msize;  // rank(l) + rank(r): size of the capacitance M
  int mmax = 0;
};

// Factored inverse of the compressed operator in precision F:
//   leaf inverses, Q_c = A_node^{-1} (U S)_c for every depth, and the
//   capacitance inverses M^{-1} = (I + K C)^{-1} of every internal node.
// An instantiated HIP graph; moving leaves the source empty.
struct GraphExec {
  hipGraph_t g = nullptr;
  hipGraphExec_t e = nullptr;
  GraphExec() = default;
  GraphExec(const GraphExec&) = delete;
  GraphExec& operator=(const GraphExec&) = delete;
  GraphExec(GraphExec&& o) noexcept : g(o.g), e(o.e) { o.g = nullptr; o.e = nullptr; }
  GraphExec& operator=(GraphExec&& o) noexcept {
    if (this != &o) {
      reset();
      g = o.g;
      e = o.e;
      o.g = nullptr;
      o.e = nullptr;
    }
    return *this;
  }
  ~GraphExec() { reset(); }
  void reset() {
    if (e) (void)hipGraphExecDestroy(e);
    if (g) (void)hipGraphDestroy(g);
    e = nullptr;
    g = nullptr;
  }
};

// Solves of one column count launch eagerly kEagerSolves times after a
// factorisation before they are recorded into a graph: recording costs
// milliseconds, which one or two solves per factorisation never win back.
constexpr int kEagerSolves = 2;

// The apply-inverse of one factorisation for a fixed column count: every
// launch planned (descriptors on the device, buffers sized) and recorded
// into one HIP graph that reads fa.b and writes fa.x.

template <class F>
struct SolvePlan {
  int ncols = 0;
  std::map<int, int> uses;  // eager applies per column count since the factorisation
  kern::NNPlan leaf;
  std::vector<kern::TNPlan> tn;
  std::vector<kern::NNPlan> mn, up;
  std::vector<int64_t> mm;  // per parent depth: capacitance leading dimension
  DevBuf<F> part;
  // FP64: every parent depth's capacitance solves as node tasks
  DevBuf<kern::BGemm> cap_ops;
  DevBuf<kern::NodeTask> cap_tasks;
  std::vector<int> cap_first;  // per parent depth: its first task
  GraphExec graph;
  // device addresses of the factorisation's buffers that the graph reads
  // and writes, as recorded: eager applies with more columns may reallocate
  // (grow) any of them, and the graph is then recorded again
  std::vector<const void*> bufs;
  void clear() {
    graph.reset();
    tn.clear();
    mn.clear();
    up.clear();
    mm.clear();
    cap_first.clear();
    part.release();
    bufs.clear();
    ncols = 0;
  }
};

// Leaf and capacitance solves of the FP64 factorisation apply the explicit
// inverse and then one step of refinement against the block itself,
// x = X b, x += X (b - M x): backward stable like an LU solve (fixed
// precision refinement, Higham 2002, sec. 12.3) while every step stays a
// batched GEMM.  With the inverse alone the recursive Woodbury solve loses
// about two digits against the reference's LU-based solve on the same
// operator (numpy emulation of this factorisation on the sigma = 0.05,
// lam = 1e-2 Gaussian system, N = 4096: 9.7e-11 with inverses, 1.9e-12
// with the refinement step, 1.5e-12 with LU solves, 1.2e-12 for the
// reference algorithm).  The FP32 factorisation of the mixed path leaves
// accuracy to the FP64 refinement around it.
template <class F>
constexpr bool kRefineBlocks = std::is_same<F, double>::value;

template <class F>
struct HFactor {
  DevBuf<F> Dinv;                 // nleaf x lmax x lmax
  DevBuf<F> Wall;                 // N x Rtot
  std::vector<DevBuf<F>> Vown;    // per depth, FP32 copies of V
  std::vector<const F*> V;        // per depth
  std::vector<DevBuf<F>> Minv;    // per parent depth: np x mmax x mmax
  // FP64: the capacitance matrices themselves, for the refinement step of
  // every capacitance solve (kRefineBlocks)
  std::vector<DevBuf<F>> Mmat;
  DevBuf<F> tvec, coef, wtmp, b, x;
  DevBuf<F> rtmp;                 // leaf-solve residual (N x columns)
  SolvePlan<F> plan;
  kern::ApPlan<F> ap;             // the fused apply-inverse (one persistent launch)
  bool ready = false;
  // invalidate; the device arrays are kept for the next factorisation
  void clear() {
    plan.clear();
    plan.uses.clear();
    ap.clear();
    V.clear();
    ready = false;
  }
};

class HodlrEngine {
 public:
  explicit HodlrEngine(const hmat_options& o);
  ~HodlrEngine();

  void build(const GenDesc& g, int64_t N, int leaf, const double* host_omega, int64_t omega_len);
  // Independent systems of the given sizes on consecutive rows (a forest):
  // one operator, block diagonal across systems, built and solved level by
  // level for all systems at once.
  void build_forest(const GenDesc& g, const std::vector<int64_t>& sizes, int leaf);
  void matvec(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy);
  // the same without the final synchronisation
  void matvec_enqueue(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy);
  // The fused matvec with its partial products in caller-provided device
  // scratch of `cap` doubles (groups of up to four columns, one after
  // another), then a synchronisation: no allocation, copy or descriptor
  // upload.  matvec_scratch_need(nrhs) doubles are needed; HMAT_ERR_ARG when
  // cap is smaller or the operator's leaves exceed the fused kernels' LDS.
  size_t matvec_scratch_need(int nrhs) const;
  void matvec_with_scratch(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy, double* scratch,
                           size_t cap);
  int device = 0;  // the HIP device current when the engine was created
  void factor();
  void solve(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);
  void to_dense_host(double* A, int64_t lda);
  // Operator file: options, tree sizes, per-depth ranks, (U S) and V, leaf
  // blocks.  load() replaces the handle's operator and options.
  void save(const std::string& path);
  void load(const std::string& path);
  // A_H^T x and A_H^{-1 T} b (adjoints): the transposed operator is a
  // HODLR matrix with (U S) and V exchanged, ranks taken from the sibling
  // and the leaf blocks transposed.  Its factorisation is kept separately.
  void matvec_transpose(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy);
  void solve_transpose(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);

  // Build + factor + solve of independent systems (a forest, as
  // build_forest) in one call, X = A^{-1} B for ncols columns.  With
  // use_graph (fixed rank, FP64 only) the work follows a fixed schedule
  // without host synchronisation (shifted CholeskyQR3, a refinement count
  // fixed when the batch shape is first seen) and is replayed from one HIP
  // graph: the first call for a shape runs that schedule eagerly while its
  // descriptors are kept on the device, and records the same launches into
  // the graph; a later call with the same shape (system sizes, leaf size,
  // kernel kind, point arrays and width p0, column count, options) replays
  // it.  New point values written in place and a new diagonal shift p1 (a
  // device scalar added to the leaf blocks) replay the same graph.  One
  // download after each replay checks it (Cholesky and pivot statuses, the
  // residual of every column); a replay whose residual misses the target is
  // refined further eagerly, one whose statuses fail is redone eagerly and
  // recorded again.
  // Any other build, load or transposed operation drops the graph.
  enum BatchPath { kBatchEager = 0, kBatchRecorded = 1, kBatchReplayed = 2, kBatchRedone = 3 };
  int batch_solve(const GenDesc& g, const std::vector<int64_t>& sizes, int leaf, const double* dB, int64_t ldb,
                  int ncols, double* dX, int64_t ldx, bool use_graph);
  // why the last batch_solve with use_graph did not record or replay (empty when it did)
  std::string batch_note;

  static std::vector<std::pair<int64_t, int>> omega_schedule(const hmat_options& o, int64_t N,
                                                             int leaf);

  hmat_options opt;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int64_t N = 0;
  int leaf = 0;
  Tree tree;
  std::vector<HLevel> lv;  // index = depth (lv[0] unused)
  DevBuf<double> D;        // nleaf x lmax x lmax
  bool built = false, factored = false;
  std::vector<HParents> par;  // index = depth of the parent
  std::vector<int64_t> colofs;
  int64_t Rtot = 0;
  // FP64 refinement: on by default in mixed precision, opt-in in FP64
  double refine_rtol = 0.0;
  int refine_max = 30;
  bool solve_graph = true;  // replay the apply-inverse from a captured HIP graph
  // apply-inverse as one persistent launch per four columns: 0 never,
  // 1 where it is the faster path, 2 whenever the device holds its grid
  int fused_solve = 1;
  double group_bytes = kGroupScratchBytes;  // scratch budget of one group of depths
  int last_iters = 0;
  double last_resid = -1.0;

 private:
  void reset_for_build(int leaf);
  void build_levels(const GenDesc& g, const double* host_omega, int64_t omega_len);
  void build_group(int c0, int c1, const GenDesc& g, const std::vector<int64_t>& omega_off,
                   const double* host_omega);
  int group_kcap(int c0, int c1) const;
  template <class F> void factor_t(HFactor<F>& fa);
  // cap / capops: the depth's capacitance solves as prebuilt node tasks
  // (FP64 applies); null: three segmented products
  template <class F> void correction_t(HFactor<F>& fa, int j, F* W, int64_t ldw, int ncols,
                                       const kern::NodeTask* cap = nullptr, const kern::BGemm* capops = nullptr);
  template <class F> void apply_inverse_t(HFactor<F>& fa, const F* B, int64_t ldb, int ncols, F* X,
                                          int64_t ldx);
  template <class F> void leaf_apply_t(const F* Dm, const F* B, int64_t ldb, int ncols, F* Y,
                                       int64_t ldy, F alpha = F(1), F beta = F(0));
  // Y = D^{-1} B over all leaves (inverse, then in FP64 one refinement step
  // whose residual goes to R, ld N; R may be B itself when B is scratch)
  template <class F> void leaf_solve_t(HFactor<F>& fa, const F* B, int64_t ldb, int ncols, F* Y, int64_t ldy,
                                       F* R);
  // graph-captured apply-inverse (fa.b -> fa.x, N x ncols, ld N)
  template <class F> void plan_solve(HFactor<F>& fa, int ncols);
  // descriptors of the fused apply-inverse (after a factorisation)
  template <class F> void plan_fused(HFactor<F>& fa);
  template <class F> bool fused_ready(const HFactor<F>& fa, int ncols) const;
  // raise the fused kernel's error flag as an exception (after a sync)
  template <class F> void check_fused(HFactor<F>& fa);
  void check_fused_active();
  template <class F> void run_solve_graph(HFactor<F>& fa, int ncols);
  void apply_inverse(const double* B, int64_t ldb, int ncols, double* X, int64_t ldx);
  void apply_inverse_impl(const double* B, int64_t ldb, int ncols, double* X, int64_t ldx);
  double solve_bytes(int ncols) const;
  // fused matvec plan (built once per build) and its launch without a host sync
  void prepare_matvec();
  void flip();  // switch between A_H and A_H^T in place (involution)
  bool transposed_ = false, factoredT_ = false;
  HFactor<double> f64T_;
  HFactor<float> f32T_;
  void matvec_async(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy);
  double matvec_bytes(int nrhs) const;

  HFactor<double> f64_;
  HFactor<float> f32_;
  DevBuf<double> ir_y_, ir_r_, ir_n_, ir_p_, ir_best_;

  kern::Workspace ws_;
  Orthogonalizer orth_;
  DevBuf<SegNN> dnn_;
  DevBuf<SmallMat> dsm_;
  DevBuf<BlockDesc> dbd_;
  DevBuf<double> Yb_, Xb_, Tb_, Rb_, Vr_, Sv_, Wk_, TU_, TV_;
  DevBuf<double> tvec_;
  DevBuf<int> info_;
  kern::MvPlan mv_;
  DevBuf<kern::MvLeaf> mv_leaves_;
  DevBuf<kern::MvLevel> mv_lvls_;
  DevBuf<kern::MvEntry> mv_ent_;
  DevBuf<kern::MvSeg> mv_segs_;
  DevBuf<double> mv_part_, mv_t_;
  int mv_width_ = 0;  // widest fused matvec pass (16, 8 or 4 columns; 0: none fits LDS)
  double mv_bytes_base_ = 0.0;  // D + (U S) + V bytes of one pass
  int last_ninfo_ = 0;          // pivot statuses written by the last factorisation

  // refinement of X toward A_H^{-1} B (the loop of solve() after the apply)
  void refine_loop(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx);
  // Fixed-schedule state of a recorded build: no host synchronisation or
  // host decision (ranks fixed, statuses kept for a later check), the leaf
  // blocks' diagonal shift read from a device scalar.
  bool fixed_ = false, capturing_ = false;
  const double* diag_shift_ = nullptr;
  // X = A^{-1} B in the staged batch buffers with `iters` refinement steps;
  // adaptive: as many as the target needs (host checks), `iters` receives the count
  void fixed_solve(int ncols, int& iters, bool adaptive, double& resid);
  // statuses and residual after a replay: "" when the statuses are good
  std::string batch_check(int ncols, double& resid);
  // every device buffer a recorded batch graph reads or writes: a replay
  // needs each at the address it had when the graph was recorded
  std::vector<const void*> batch_buffers() const;
  struct BatchGraph {
    std::vector<int64_t> sizes;
    int leaf = 0, ncols = 0, kind = -1;
    const double *x = nullptr, *y = nullptr;
    double p0 = 0.0;
    hmat_options opt{};
    int iters = 0, ninfo = 0;
    size_t orth_calls = 0;
    std::vector<const void*> bufs;  // batch_buffers() when recorded
    bool valid = false;
    DevBuf<double> b, x_, shift;  // staged right-hand sides and solution, diagonal shift
    std::unique_ptr<StageLog> log;
    GraphExec graph;
    void drop() {
      graph.reset();
      log.reset();
      valid = false;
    }
  } bg_;
};

// Allocation-free matvec (noalloc.hip): the fused kernels with their
// partials in a fixed per-device scratch array of
// hodlr_matvec_noalloc_capacity() doubles.
void hodlr_matvec_noalloc(HodlrEngine& e, const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy);
size_t hodlr_matvec_noalloc_capacity();

}  // namespace hmat
