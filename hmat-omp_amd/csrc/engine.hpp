// Tree topology and shared level-batched building blocks.
#pragma once
#include <cmath>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace hmat {

// Binary cluster tree: mid = (s+e)/2, leaf when len <= leaf_size
// (hodlr_routines_omp.h:36-43, hss_routines_omp.h:361-374).
struct TNode {
  int64_t start = 0, len = 0;
  int depth = 0, parent = -1, left = -1, right = -1;
  bool leaf = false;
  int pos = 0;       // index within its depth's list
  int leaf_idx = -1;
  int ref_id = 0;    // HSSNode::id of the reference (both children numbered before recursing)
};

struct Tree {
  std::vector<TNode> nodes;               // nodes[0] = root
  std::vector<std::vector<int>> by_depth; // node ids per depth, in start order
  std::vector<int> leaves;                // in start order
  int max_depth = 0;
  int64_t lmax = 0;

  void build(int64_t N, int64_t leaf) {
    nodes.clear();
    by_depth.clear();
    leaves.clear();
    nodes.push_back(TNode{});
    nodes[0].start = 0;
    nodes[0].len = N;
    int next_ref = 1;
    grow(0, leaf, next_ref);
    max_depth = 0;
    for (auto& n : nodes) max_depth = std::max(max_depth, n.depth);
    by_depth.assign(max_depth + 1, {});
    // depth lists in start order: a pre-order walk visits each depth left to right
    walk(0);
    lmax = 0;
    for (int id : leaves) lmax = std::max(lmax, nodes[id].len);
  }

  // Forest of independent systems (batched solves): system s owns rows
  // [sum_{t<s} sizes[t], ... + sizes[s]) and is bisected on its own; the
  // roots are the depth-0 nodes, so every depth holds the nodes of all
  // systems and a level still takes one launch.
  void build_forest(const std::vector<int64_t>& sizes, int64_t leaf) {
    nodes.clear();
    by_depth.clear();
    leaves.clear();
    int next_ref = 0;
    std::vector<int> roots;
    int64_t off = 0;
    for (int64_t n : sizes) {
      TNode r;
      r.start = off;
      r.len = n;
      r.ref_id = next_ref++;
      roots.push_back((int)nodes.size());
      nodes.push_back(r);
      off += n;
    }
    for (int id : roots) grow(id, leaf, next_ref);
    max_depth = 0;
    for (auto& n : nodes) max_depth = std::max(max_depth, n.depth);
    by_depth.assign(max_depth + 1, {});
    for (int id : roots) walk(id);
    lmax = 0;
    for (int id : leaves) lmax = std::max(lmax, nodes[id].len);
  }

 private:
  void grow(int id, int64_t leaf, int& next_ref) {
    TNode& n = nodes[id];
    if (n.len <= leaf) {
      nodes[id].leaf = true;
      return;
    }
    const int64_t s = n.start, e = n.start + n.len, mid = (s + e) / 2;
    const int d = n.depth;
    TNode l, r;
    l.start = s; l.len = mid - s; l.depth = d + 1; l.parent = id; l.ref_id = next_ref++;
    r.start = mid; r.len = e - mid; r.depth = d + 1; r.parent = id; r.ref_id = next_ref++;
    const int li = (int)nodes.size();
    nodes.push_back(l);
    nodes.push_back(r);
    nodes[id].left = li;
    nodes[id].right = li + 1;
    grow(li, leaf, next_ref);
    grow(li + 1, leaf, next_ref);
  }
  void walk(int id) {
    TNode& n = nodes[id];
    n.pos = (int)by_depth[n.depth].size();
    by_depth[n.depth].push_back(id);
    if (n.leaf) {
      n.leaf_idx = (int)leaves.size();
      leaves.push_back(id);
      return;
    }
    walk(n.left);
    walk(n.right);
  }
};

inline GenDesc gen_from_kernel(const hmat_kernel& k, const double* dx, const double* dy) {
  GenDesc g;
  g.x = dx;
  g.y = dy;
  g.p0 = k.p0;
  g.p1 = k.p1;
  switch (k.type) {
    case HMAT_KERNEL_GAUSSIAN: g.kind = GenKind::Gaussian; break;
    case HMAT_KERNEL_CAUCHY: g.kind = GenKind::Cauchy; break;
    case HMAT_KERNEL_REF_CAUCHY: g.kind = GenKind::RefCauchy; break;
    case HMAT_KERNEL_REF_GAUSSIAN: g.kind = GenKind::RefGaussian; break;
    case HMAT_KERNEL_REF_COULOMB: g.kind = GenKind::RefCoulomb; break;
    case HMAT_KERNEL_GAUSSIAN_DSIGMA: g.kind = GenKind::GaussianDSigma; break;
    default: throw Error(HMAT_ERR_ARG, "unknown kernel type");
  }
  return g;
}

// Segmented shifted CholeskyQR (Fukaya et al., SIAM J. Sci. Comput. 2020):
// one shifted pass, then unshifted passes until every segment's Gram matrix
// is the identity to kOrthTol = 1e-12 (two passes and the check for
// well-conditioned input; sketches of numerically low-rank blocks, cond
// ~1e16, take more).  Q overwrites `Y`; optionally R = R_p ... R_2 R_1.
constexpr double kOrthTol = 1e-12;
constexpr int kOrthMaxPasses = 8;
struct OrthSeg {
  int64_t off;  // row offset of the segment inside the N x kcap arrays
  int32_t len, k;
};

class Orthogonalizer {
 public:
  // Y, T: device arrays with leading dimension ld, at least kcap columns.
  // Rout (optional): nseg matrices of kcap x kcap (ld kcap).
  void run(const std::vector<OrthSeg>& segs, double* Y, double* T, int64_t ld, int kcap,
           double* Rout, hipStream_t s);
  // One shifted pass and nothing else (no check, no synchronisation): a
  // basis of the same span with orthogonality defect ~u cond(Y)^2, for
  // intermediate bases whose input is well conditioned (cond(Y)^2 << 1/u).
  // A breakdown of its Cholesky factorisation (non-finite input) is
  // reported by the next run() or check_once().
  void run_once(const std::vector<OrthSeg>& segs, double* Y, double* T, int64_t ld, int kcap,
                hipStream_t s);
  // Download and check the statuses of the run_once passes so far.
  void check_once(hipStream_t s);
  // Schedule for builds recorded into a graph: run() becomes one shifted
  // pass, then unshifted passes (an unshifted factor that breaks down is
  // redone shifted in the same launch), each preceded by the Gram matrix
  // and Cholesky factor that also check the current basis, and a final
  // check.  With `choose`, the k-th call since set_fixed(true, ...) takes
  // passes until its basis is orthonormal to kOrthTol (downloads; at most
  // kOrthMaxPasses) and records their number; without, it takes the number
  // recorded for the k-th call, with no download and no host decision.  The
  // staged descriptors do not depend on that number.  Each call's statuses
  // and final Gram deviations go to buffers of their own, so a replay of
  // the recorded launches writes them again in place.
  void set_fixed(bool on, bool choose = false) {
    fixed_ = on;
    fixed_choose_ = choose;
    if (on) fixed_call_ = 0;
  }
  bool fixed() const { return fixed_; }
  // every device buffer the calls use (a recorded graph holds these addresses)
  void addresses(std::vector<const void*>& v) const {
    v.insert(v.end(), {G_.get(), R_.get(), Rchk_.get(), Racc_.get(), Rtmp_.get(), gerr_.get(), info_.get(),
                       once_info_.get(), chk_.get(), dnn_.get(), dsm_.get(), plan_.blob.get()});
    for (const auto& b : fixed_info_) v.push_back(b.get());
    for (const auto& b : fixed_gerr_) v.push_back(b.get());
    ws_.addresses(v);
  }
  size_t fixed_calls() const { return fixed_call_; }
  // The statuses of the first `ncalls` fixed-schedule calls since the last
  // set_fixed(true): empty when every Gram matrix was finite, every shifted
  // factor succeeded and every final basis is orthonormal to
  // max(kOrthTol, 10 x the deviation recorded for the call), else what failed.
  std::string check_fixed(size_t ncalls, hipStream_t s);

 private:
  void run_fixed(const std::vector<OrthSeg>& segs, double* Y, double* T, int64_t ld, int kcap, double* Rout,
                 hipStream_t s);
  bool fixed_ = false, fixed_choose_ = false;
  size_t fixed_call_ = 0;
  std::vector<DevBuf<int>> fixed_info_;      // per call: shifted-pass statuses, then the final check's
  std::vector<DevBuf<double>> fixed_gerr_;   // per call: the final check's Gram deviations
  std::vector<size_t> fixed_nb_;              // per call: nseg
  std::vector<int> fixed_passes_;             // per call: unshifted passes recorded
  std::vector<double> fixed_dev_;             // per call: final deviation when recorded
  // the launches of one call, planned once: Gram products, Cholesky
  // descriptors and shifts, triangular solves, R products
  struct Plan {
    double panel = 0;      // sum of len * k over the segments (elements of the N x k panel)
    DevBuf<uint8_t> blob;  // every descriptor array below, uploaded with one copy
    const SegTN* tn_segs = nullptr;
    const kern::TNItem* tn_items = nullptr;
    const kern::TNOut* tn_outs = nullptr;
    int nitems = 0, nouts = 0, nslots = 0;
    const SmallMat* mats = nullptr;
    const double* shift_on = nullptr;
    const double* shift_off = nullptr;
    const SegNN* tr_segs = nullptr;
    const kern::NNTile* tr_tiles = nullptr;
    int tr_ntiles = 0, kmax = 0;
    const SegNN* acc_segs = nullptr;
    const kern::NNTile* acc_tiles = nullptr;
    int acc_ntiles = 0, acc_mode = 0;
    int nmax = 0;
  } plan_;
  void make_plan(const std::vector<OrthSeg>& segs, int kcap, bool with_acc, hipStream_t s);
  void gram(const double* Y, int64_t ld, int kcap, hipStream_t s);
  void trsm(const double* X, int64_t ld, const double* R, int kcap, double* Y, hipStream_t s);
  void acc_product(int kcap, hipStream_t s);
  kern::Workspace ws_;
  DevBuf<double> G_, R_, Rchk_, Racc_, Rtmp_, gerr_;  // Rchk_: the check pass's R (discarded)
  DevBuf<int> info_;
  DevBuf<int> once_info_;  // run_once statuses not yet checked: [0, once_off_)
  size_t once_off_ = 0;
  DevBuf<uint8_t> chk_;  // both passes' info and Gram deviations, downloaded together
  DevBuf<SegNN> dnn_;
  DevBuf<SmallMat> dsm_;
  // One CholeskyQR pass in -> out.  Returns false (and leaves out alone)
  // when `in` is already orthonormal to kOrthTol and `shifted` is false.
  bool pass(const std::vector<OrthSeg>& segs, const double* in, double* out, int64_t ld, int kcap,
            bool shifted, hipStream_t s);
};

}  // namespace hmat
