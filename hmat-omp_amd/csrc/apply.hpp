// Fused apply-inverse of the HODLR factorisation (one persistent launch).
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

#include "common.hpp"

namespace hmat {
namespace kern {

struct ApLeaf {
  int64_t start;
  int32_t len, pad_;
};
// One parent depth j (children at depth j + 1).
template <class F>
struct ApLevel {
  const F* V;     // V of depth j + 1, N x rmax, ld N
  const F* Q;     // Q of depth j + 1 (A_child^{-1} (U S)_child), N x rmax, ld N
  const F* Minv;  // np x mm x mm capacitance inverses of depth j
  const F* M;     // the capacitance matrices themselves (refined solves), or null
  int32_t mm, pad_;
};
// (leaf, parent depth j): the depth-(j+1) node s holding the leaf and its parent p.
struct ApEntry {
  int32_t p;      // global parent index, -1 when the leaf is shallower than depth j + 1
  int32_t vcols;  // columns of V at s
  int32_t qcols;  // columns of Q at s
  int32_t coff;   // first coefficient of the parent's solution that multiplies Q_s
};
// An internal node: its leaves are [leaf0, leaf1), the left child's [leaf0, lsplit).
struct ApParent {
  int64_t moff;   // offset of its capacitance matrix in the depth's array
  int32_t leaf0, lsplit, leaf1;
  int32_t rr, rl; // rows of the capacitance right-hand side from the left child (V_l^T x_l: rr) and right (rl)
  int32_t level, pad_[2];  // explicit padding: every byte of an uploaded descriptor is defined
};
static_assert(sizeof(ApParent) == 40, "ApParent has no implicit padding");

template <class F>
struct ApArgs {
  const ApLeaf* leaves;
  const ApEntry* ent;  // nleaf x L
  const ApLevel<F>* lv;
  const ApParent* par;
  const F* Dinv;
  const F* D;  // leaf blocks for the refinement step, or null
  int64_t lm;
  const F* B;
  int64_t ldb;
  F* X;
  int64_t ldx;
  int64_t ldv;  // leading dimension of V and Q (N)
  F* part;   // nleaf x KM x NR
  F* coef;   // npar x MM x NR
  int* cnt;  // npar arrival counters
  int* flag; // npar "coefficients ready" flags
  int* err;  // set when a wait exceeded its spin limit
  int nleaf, L, KM, MM;
  int64_t spin_limit;
};

// Device memory that bypasses the caches (hipDeviceMallocUncached), for
// the fused kernel's traffic between workgroups: partials, coefficients,
// counters and flags written by one workgroup and read by others.  On the
// MI355X each XCD has its own L2, so through cached memory every such
// hand-over needs an L2 write-back by the writer and an invalidation by
// the reader; uncached stores and loads go to memory and need neither.
// Blocks are never freed before the process ends (a released block goes
// back to a process-wide pool for the next plan): freeing uncached memory
// after a kernel read it is not safe on this platform.
class UcBuf {
 public:
  UcBuf() = default;
  UcBuf(const UcBuf&) = delete;
  UcBuf& operator=(const UcBuf&) = delete;
  UcBuf(UcBuf&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  UcBuf& operator=(UcBuf&& o) noexcept {
    std::swap(p_, o.p_);  // the moved-from object releases what this held
    std::swap(n_, o.n_);
    return *this;
  }
  ~UcBuf();
  void reserve(size_t bytes);  // grow-only
  template <class T>
  T* as() const { return static_cast<T*>(p_); }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// Everything the launch needs beyond the per-call buffers: capacity of the
// device for the persistent grid, descriptor arrays on the device.
template <class F>
struct ApPlan {
  DevBuf<ApLeaf> leaves;
  DevBuf<ApEntry> ent;
  DevBuf<ApLevel<F>> lv;
  DevBuf<ApParent> par;
  UcBuf part, coef;  // F: nleaf x KM x 4, npar x MM x 4
  UcBuf sync;        // int: cnt (npar), flag (npar), err (1)
  int nleaf = 0, L = 0, KM = 0, MM = 0, npar = 0;
  int64_t lm = 0;
  bool ok = false;
  void clear() { ok = false; }
  void addresses(std::vector<const void*>& v) const {
    v.insert(v.end(), {leaves.get(), ent.get(), lv.get(), par.get(), part.as<void>(), coef.as<void>(),
                       sync.as<void>()});
  }
};

// X = A_H^{-1} B for ncols columns (groups of up to four per launch): the
// leaf solves, then every depth's capacitance correction, deepest first,
// in one cooperative launch per column group.  Returns false (and does
// nothing) when the device cannot hold the persistent grid.
template <class F>
bool apply_inverse_fused(const ApPlan<F>& p, const F* Dinv, const F* D, const F* B, int64_t ldb, int ncols, F* X,
                         int64_t ldx, int64_t ldv, hipStream_t s);
// The largest grid of the fused kernel that the device holds at once for
// this plan (0: the kernel cannot run here, use the multi-launch path).
template <class F>
int apply_fused_grid(const ApPlan<F>& p, bool refine, int nr);
// Spin limit of every wait of the fused kernel (tests lower it to check the
// error path; a wait that exceeds it raises the error flag).
void apply_fused_set_spin_limit(int64_t n);

}  // namespace kern
}  // namespace hmat
