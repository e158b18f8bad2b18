// One-launch HSS matvec for up to 16 right-hand sides (persistent,
// cooperative).
#pragma once
#include <cstdint>

#include "apply.hpp"
#include "common.hpp"

namespace hmat {
namespace kern {

struct HmvLeaf {
  int64_t start;
  int32_t len, id;  // rows, tree node id
  int32_t rU, rV;
};
// One node of a sweep.  Up (kind 0): g_p = W^T [g_a; g_b] with W at o1
// ((ra + rb) x rp, ld ld1).  Down (kind 1): f_a = B_ab g_b + R(0:ra, :) f_p,
// f_b = B_ba g_a + R(ra:ra+rb, :) f_p with B_ab at o1 (ra x gb), B_ba at o2
// (rb x ga), both ld ld1, and R at o3 ((ra + rb) x rp, ld ld2); rp = 0 at
// the root (no f_p).
struct HmvTask {
  int32_t kind, p, a, b;
  int32_t ra, rb, rp, ga, gb, ld1, ld2, pad_;
  int64_t o1, o2, o3;
};
static_assert(sizeof(HmvTask) == 72, "HmvTask has no implicit padding");

struct HmvArgs {
  const HmvLeaf* leaves;
  const HmvTask* tasks;
  const double* D;  // leaf blocks, nleaf x lm x lm
  const double* U;  // leaf bases, N x rmax, ld N
  const double* V;
  const double* pool;
  const double* X;
  double* Y;
  double* G;  // per node g (up-sweep), nn x ldv x NR, uncached
  double* F;  // per node f (down-sweep), nn x ldv x NR, uncached
  int* gflag;
  int* fflag;
  int* counter;  // next task
  int* err;      // a wait exceeded its spin limit
  int64_t N, lm, spin_limit, ldx, ldy;
  int nleaf, ntasks, ldv, pad_;
};

// Device state of the one-launch matvec of one operator: descriptors, the
// per-node vectors and flags.  Built once per operator (build or load).
struct HmvPlan {
  DevBuf<HmvLeaf> leaves;
  DevBuf<HmvTask> tasks;
  UcBuf vec;   // G then F, 2 x nn x ldv x 16 doubles
  UcBuf sync;  // gflag (nn), fflag (nn), counter, err
  int* err_host = nullptr;  // pinned copy of err after each call
  int nleaf = 0, ntasks = 0, nn = 0, ldv = 0;
  int grid[5] = {0, 0, 0, 0, 0};  // hss_mv_grid for 1, 2, 4, 8, 16 columns
  int64_t lm = 0;
  bool ok = false;
  HmvPlan() = default;
  HmvPlan(const HmvPlan&) = delete;
  HmvPlan& operator=(const HmvPlan&) = delete;
  ~HmvPlan();
};

// Largest grid the device holds at once for this plan and 2^l columns per
// launch (0: not available).
int hss_mv_grid(const HmvPlan& p, int l);
// Y = A_HSS X for ncols columns (X, Y: device, column-major), enqueued on
// s, up to 16 columns per launch.  Returns false when a launch was refused
// (the columns before it are done).
bool hss_mv_launch(HmvPlan& p, const double* D, const double* U, const double* V, const double* pool, int64_t N,
                   const double* X, int64_t ldx, int ncols, double* Y, int64_t ldy, hipStream_t s);
// Waits for the launch and throws when one of its waits exceeded its spin
// limit (the error flag is cleared by the next launch).
void hss_mv_finish(HmvPlan& p, hipStream_t s);
// Spin limit of the waits (tests lower it to reach the error path).
void hss_mv_set_spin_limit(int64_t n);

}  // namespace kern
}  // namespace hmat
