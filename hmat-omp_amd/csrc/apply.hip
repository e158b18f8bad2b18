// Fused apply-inverse of the HODLR factorisation: one persistent launch per
// group of up to four right-hand sides.
//
// Reference: HODLRMatrixOMP::solve_woodbury (include/hodlr_routines_omp.h:
// 147-240), restated factor-once as X = D^{-1} B followed by one correction
// per parent depth, deepest first:
//   t_p  = [V_l^T X_l ; V_r^T X_r]      (rows of the two children of p)
//   c_p  = M_p^{-1} t_p                  (refined once against M_p in FP64)
//   X_l -= Q_l c_p[rr:],  X_r -= Q_r c_p[:rr]
// The multi-launch form takes three launches per depth plus the leaf
// solves, and each launch drains the GPU before the next one starts.  Here
// every workgroup owns leaves (grid-stride over the leaves, the grid no
// larger than the device holds at once, launched cooperatively so that
// every workgroup is resident):
//   1. each owned leaf: X = D^{-1} B over its rows (and in FP64 one
//      refinement step against D), B and the intermediate vectors staged in
//      LDS, the leaf blocks read coalesced down their columns;
//   2. per parent depth, deepest first: each owned leaf writes its rows'
//      share of V^T X to its slot of `part` and counts itself in at its
//      parent; the last leaf to arrive sums the parent's partials in leaf
//      order (so the result does not depend on arrival order), solves the
//      capacitance system and raises the parent's flag; every leaf then
//      waits for its parent's flag and applies X -= Q c over its rows.
// V and Q are read once per launch, coalesced; X moves between HBM and the
// workgroup only in leaf-sized pieces.  Every wait is bounded: a wait that
// exceeds its spin limit raises `err` and its workgroup returns, so every
// wave ends even if a co-resident grid were not obtained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <utility>
#include <vector>

#include "apply.hpp"

namespace hmat {
namespace kern {

constexpr int kApWG = 256;
constexpr int kApWaves = kApWG / 64;
constexpr int kApMaxRpt = 4;  // rows per thread: leaves of up to 1024 rows

namespace {
std::atomic<int64_t> g_spin_limit{(int64_t)1 << 22};
std::mutex g_uc_mu;
std::vector<std::pair<void*, size_t>>& uc_pool() {  // released blocks, never freed
  static std::vector<std::pair<void*, size_t>> v;
  return v;
}
}  // namespace
void apply_fused_set_spin_limit(int64_t n) { g_spin_limit = n > 0 ? n : ((int64_t)1 << 22); }

UcBuf::~UcBuf() {
  if (!p_) return;
  std::lock_guard<std::mutex> lk(g_uc_mu);
  uc_pool().push_back({p_, n_});
}

void UcBuf::reserve(size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  if (bytes <= n_) return;
  std::lock_guard<std::mutex> lk(g_uc_mu);
  if (p_) uc_pool().push_back({p_, n_});
  p_ = nullptr;
  n_ = 0;
  // the smallest pooled block that fits
  auto& pool = uc_pool();
  size_t best = pool.size();
  for (size_t i = 0; i < pool.size(); ++i)
    if (pool[i].second >= bytes && (best == pool.size() || pool[i].second < pool[best].second)) best = i;
  if (best < pool.size()) {
    p_ = pool[best].first;
    n_ = pool[best].second;
    pool.erase(pool.begin() + (std::ptrdiff_t)best);
    return;
  }
  HIP_CHECK(hipExtMallocWithFlags(&p_, bytes, hipDeviceMallocUncached));
  n_ = bytes;
}

template <class F>
using gptr = const __attribute__((address_space(1))) F*;

template <int S, class F>
__device__ __forceinline__ void ap_halving(F (&v)[16], int lane) {
  if constexpr (S >= 1) {
    const bool upper = (lane & S) != 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const F send = upper ? v[i] : v[i + S];
      const F keep = upper ? v[i + S] : v[i];
      v[i] = keep + __shfl_xor(send, S);
    }
    ap_halving<S / 2>(v, lane);
  }
}
// 16 values per lane in; lane l ends with the wave total of v[l & 15]
template <class F>
__device__ __forceinline__ F ap_reduce16(F (&v)[16]) {
  const int lane = threadIdx.x & 63;
  ap_halving<8>(v, lane);
  F s = v[0];
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s;
}

// Hand-over between workgroups.  Everything one workgroup writes for
// another (partials, coefficients, counters, flags) is written and read
// with relaxed agent-scope atomic stores and loads, which are coherent
// across the XCDs by themselves (no L2 write-back or invalidation, unlike
// a release or acquire fence, which on this GPU writes back or invalidates
// the whole L2 of the XCD).  The order "data, then signal" comes from the
// writer waiting for its stores to complete (s_waitcnt vmcnt(0), which on
// this architecture counts stores too) before the barrier that precedes
// the signal; the reader issues its data loads after it observed the
// signal and passed a barrier.
template <class F>
__device__ __forceinline__ void st_dev(F* p, F v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class F>
__device__ __forceinline__ F ld_dev(F* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stores_complete() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// acc[q] += sum_{j < n} M(row, j) v[j * NR + q], column j of M at M + j * ldm
template <class F, int NR>
__device__ __forceinline__ void ap_dot(gptr<F> M, int64_t ldm, const F* v, int n, F (&acc)[NR]) {
  int j = 0;
  for (; j + 8 <= n; j += 8) {
    F m[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) m[u] = M[(int64_t)(j + u) * ldm];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = fma(m[u], v[(j + u) * NR + q], acc[q]);
  }
  for (; j < n; ++j) {
    const F m = M[(int64_t)j * ldm];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = fma(m, v[j * NR + q], acc[q]);
  }
}

// Thread 0 waits for flag[p] while the workgroup waits at the barrier.
// False when the spin limit was exceeded or another workgroup raised err.
// (flag, err and the data behind them are uncached: the polling loads see
// memory, and the data loads after the barrier see what the writer
// completed before raising the flag.)
__device__ __forceinline__ bool ap_wait(int* flag, int p, int* err, int64_t limit, int* sh) {
  if (threadIdx.x == 0) {
    int ok = 1;
    int64_t it = 0;
    while (__hip_atomic_load(flag + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      if (++it > limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    *sh = ok;
  }
  __syncthreads();
  const bool ok = *sh != 0;
  __syncthreads();
  return ok;
}

// The parent's right-hand side from its leaves' partials, both children in
// one pass: dst[i] for i < T1 sums part[leaf][i] over leaves [l0, l1) (the
// left child), dst[T1 + i] for i < T2 sums part[leaf][i] over [l1, l2).
// Thread slot s of a value chunk sums the leaves a0 + s, a0 + s + S, ... in
// order (16 loads in flight), then the S slots are added in order, so the
// result does not depend on scheduling.  red: kApWG entries of LDS.
template <class F>
__device__ void ap_sum_leaves(F* part, int64_t pst, int l0, int l1, int l2, int T1, int T2, F* dst, F* red) {
  const int tid = threadIdx.x;
  const int T = T1 + T2;
  for (int i0 = 0; i0 < T; i0 += kApWG) {
    const int C = min(T - i0, kApWG);
    const int S = kApWG / C;
    const int ic = tid % C, sl = tid / C;
    const int i = i0 + ic;
    const bool left = i < T1;
    const int a1 = left ? l1 : l2;
    F acc = F(0);
    if (sl < S) {
      F* pp = part + (left ? i : i - T1);
      int l = (left ? l0 : l1) + sl;
      for (; l + 15 * S < a1; l += 16 * S) {
        F v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = ld_dev(pp + (int64_t)(l + u * S) * pst);
#pragma unroll
        for (int u = 0; u < 16; ++u) acc += v[u];
      }
      for (; l < a1; l += S) acc += ld_dev(pp + (int64_t)l * pst);
    }
    red[tid] = acc;
    __syncthreads();
    if (tid < C) {
      F sum = F(0);
      for (int u = 0; u < S; ++u) sum += red[u * C + tid];
      dst[i0 + tid] = sum;
    }
    __syncthreads();
  }
}

// y = base + sign * A v for an m x n matrix A (ld lda, global) and v in LDS
// (n x NR, row-major k * NR + q).  Thread t < S m takes row t % m and the
// k-slice t / m of S = min(4, kApWG / m) slices, so that every thread
// streams a run of one row: 32 loads in flight for one FP32 right-hand
// side, 16 otherwise (more held the FP64 kernels above 128 VGPRs).  Above
// kApWG rows: blocks of kApWG rows with one slice.  The slices are added in
// order.  y may be LDS or global; red2: 4 m NR entries of LDS.
template <class F, int NR>
__device__ void ap_matvec(gptr<F> A, int64_t lda, int m, int n, const F* v, const F* base, F sign, F* y,
                          F* red2) {
  constexpr int KB = (NR == 1 && sizeof(F) == 4) ? 32 : 16;  // loads in flight per thread
  const int tid = threadIdx.x;
  const int S = (m <= 0 || m >= kApWG) ? 1 : min(4, kApWG / m);
  const int kq = (n + S - 1) / S;
  for (int i0 = 0; i0 < m; i0 += kApWG) {
    const int rows = min(m - i0, kApWG);
    const int sl = tid / rows, i = i0 + tid % rows;
    if (sl >= S) continue;
    const int k0 = sl * kq, k1 = min(n, k0 + kq);
    F acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = F(0);
    gptr<F> Ai = A + i;
    for (int k = k0; k < k1; k += KB) {
      const int nk = min(KB, k1 - k);
      F e[KB];
#pragma unroll
      for (int u = 0; u < KB; ++u) e[u] = Ai[(int64_t)(k + min(u, nk - 1)) * lda];  // in bounds
#pragma unroll
      for (int u = 0; u < KB; ++u)
        if (u < nk)
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = fma(e[u], v[(k + u) * NR + q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) red2[((int64_t)sl * m + i) * NR + q] = acc[q];
  }
  __syncthreads();
  for (int e = tid; e < m * NR; e += kApWG) {
    F sum = F(0);
    for (int u = 0; u < S; ++u) sum += red2[(int64_t)u * m * NR + e];
    y[e] = (base ? base[e] : F(0)) + sign * sum;
  }
  __syncthreads();
}

// The last arrival at parent p: t = the parent's partials summed in leaf
// order (top rr*NR entries from the left child's leaves, the rest from the
// right child's), c = M^{-1} t (+ M^{-1}(t - M c) in FP64), c -> coef[p].
template <class F, int NR, bool REFINE>
__device__ void ap_leader(const ApArgs<F>& a, const ApLevel<F>& lv, int p, F* ts, F* cs, F* us, F* red, F* red2) {
  const ApParent P = a.par[p];
  const int msize = P.rr + P.rl;
  const int64_t pst = (int64_t)a.KM * NR;
  gptr<F> Mi = (gptr<F>)(lv.Minv + P.moff);
  gptr<F> M = REFINE ? (gptr<F>)(lv.M + P.moff) : nullptr;
  ap_sum_leaves<F>(a.part, pst, P.leaf0, P.lsplit, P.leaf1, P.rr * NR, P.rl * NR, ts, red);
  F* out = a.coef + (int64_t)p * a.MM * NR;
  F* c = cs;
  if constexpr (REFINE) {
    ap_matvec<F, NR>(Mi, lv.mm, msize, msize, ts, nullptr, F(1), cs, red2);  // c = M^{-1} t
    ap_matvec<F, NR>(M, lv.mm, msize, msize, cs, ts, F(-1), us, red2);       // r = t - M c
    ap_matvec<F, NR>(Mi, lv.mm, msize, msize, us, cs, F(1), ts, red2);       // c + M^{-1} r
    c = ts;
  } else {
    ap_matvec<F, NR>(Mi, lv.mm, msize, msize, ts, nullptr, F(1), cs, red2);
  }
  for (int e = threadIdx.x; e < msize * NR; e += kApWG) st_dev(out + e, c[e]);
}

// dst[a * NR + q] = sum over the leaf's rows r of V(r, a) X(r, q), a < vc
// (dst in LDS): 16 / NR columns of V per step, each wave's sums by shuffles,
// the four waves' sums added in order.  red: kApWaves x KMP x NR.
template <class F, int NR>
__device__ void ap_vtx(gptr<F> V, int64_t ldv, const ApLeaf& lf, int vc, const F* X, int64_t ldx, int rpt, int KMP,
                       F* red, F* dst) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  constexpr int KG = 16 / NR;
  for (int kb = 0; kb < vc; kb += KG) {
    F v[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = F(0);
    for (int t = 0; t < rpt; ++t) {
      const int r = tid + t * kApWG;
      const bool live = r < lf.len;
      const int64_t row = lf.start + (live ? r : 0);
      F x[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) x[q] = live ? X[row + (int64_t)q * ldx] : F(0);
#pragma unroll
      for (int kk = 0; kk < KG; ++kk) {
        const int k = kb + kk;
        const F vk = (live && k < vc) ? V[row + (int64_t)k * ldv] : F(0);
#pragma unroll
        for (int q = 0; q < NR; ++q) v[kk * NR + q] = fma(vk, x[q], v[kk * NR + q]);
      }
    }
    const F sum = ap_reduce16(v);  // lane & 15 = kk * NR + q
    if (lane < 16) red[w * KMP * NR + kb * NR + lane] = sum;
  }
  __syncthreads();
  for (int i = tid; i < vc * NR; i += kApWG) {
    F s = F(0);
#pragma unroll
    for (int u = 0; u < kApWaves; ++u) s += red[u * KMP * NR + i];
    dst[i] = s;
  }
  __syncthreads();
}

// Publishes a leaf's partials (src: LDS, vc * NR entries) for parent en.p
// and counts the leaf in; the last arrival sums the parent's partials,
// solves its capacitance system and raises the parent's flag.
template <class F, int NR, bool REFINE>
__device__ void ap_publish(const ApArgs<F>& a, const ApLevel<F>& lv, const ApEntry& en, int leaf, const F* src,
                           F* ts, F* cs, F* us, F* red, F* red2, int* sh) {
  const int tid = threadIdx.x;
  F* out = a.part + (int64_t)leaf * a.KM * NR;
  for (int i = tid; i < en.vcols * NR; i += kApWG) st_dev(out + i, src[i]);
  stores_complete();
  __syncthreads();  // the workgroup's partials, then its arrival
  if (tid == 0) {
    const ApParent P = a.par[en.p];
    const int old = __hip_atomic_fetch_add(a.cnt + en.p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *sh = (old == P.leaf1 - P.leaf0 - 1) ? 1 : 0;
  }
  __syncthreads();
  const bool leader = *sh != 0;
  __syncthreads();
  if (leader) {
    ap_leader<F, NR, REFINE>(a, lv, en.p, ts, cs, us, red, red2);
    stores_complete();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(a.flag + en.p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <class F, int NR, bool REFINE>
__global__ __launch_bounds__(kApWG, 4) void k_apply_inverse(ApArgs<F> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  F* sm = reinterpret_cast<F*>(smem_raw);
  const int KMP = (a.KM + 15) / 16 * 16;
  const int nred = max(kApWaves * KMP * NR, kApWG);
  F* xs = sm;                       // lm x NR
  F* rs = xs + a.lm * NR;           // lm x NR
  F* red = rs + a.lm * NR;          // nred
  F* ts = red + nred;               // MM x NR
  F* cs = ts + (int64_t)a.MM * NR;  // MM x NR
  F* us = cs + (int64_t)a.MM * NR;  // MM x NR
  F* red2 = us + (int64_t)a.MM * NR;  // waves x MM x NR
  __shared__ int sh;
  const int tid = threadIdx.x;
  const int rpt = (int)((a.lm + kApWG - 1) / kApWG);
  const int L = a.L;

  // 1. leaf solves
  for (int leaf = blockIdx.x; leaf < a.nleaf; leaf += gridDim.x) {
    const ApLeaf lf = a.leaves[leaf];
    gptr<F> Di = (gptr<F>)(a.Dinv + (int64_t)leaf * a.lm * a.lm);
    for (int e = tid; e < lf.len * NR; e += kApWG) {
      const int r = e / NR, q = e % NR;
      xs[e] = a.B[lf.start + r + (int64_t)q * a.ldb];
    }
    __syncthreads();
    F y[kApMaxRpt][NR];
    for (int t = 0; t < rpt; ++t) {
      const int r = tid + t * kApWG;
#pragma unroll
      for (int q = 0; q < NR; ++q) y[t][q] = F(0);
      if (r < lf.len) ap_dot<F, NR>(Di + r, a.lm, xs, lf.len, y[t]);
    }
    if constexpr (REFINE) {
      // rs = y; xs = b - D y; y += D^{-1} xs
      for (int t = 0; t < rpt; ++t) {
        const int r = tid + t * kApWG;
        if (r < lf.len)
#pragma unroll
          for (int q = 0; q < NR; ++q) rs[r * NR + q] = y[t][q];
      }
      __syncthreads();
      gptr<F> Dl = (gptr<F>)(a.D + (int64_t)leaf * a.lm * a.lm);
      F res[kApMaxRpt][NR];
      for (int t = 0; t < rpt; ++t) {
        const int r = tid + t * kApWG;
#pragma unroll
        for (int q = 0; q < NR; ++q) res[t][q] = F(0);
        if (r < lf.len) {
          F dy[NR];
#pragma unroll
          for (int q = 0; q < NR; ++q) dy[q] = F(0);
          ap_dot<F, NR>(Dl + r, a.lm, rs, lf.len, dy);
#pragma unroll
          for (int q = 0; q < NR; ++q) res[t][q] = xs[r * NR + q] - dy[q];
        }
      }
      __syncthreads();
      for (int t = 0; t < rpt; ++t) {
        const int r = tid + t * kApWG;
        if (r < lf.len)
#pragma unroll
          for (int q = 0; q < NR; ++q) xs[r * NR + q] = res[t][q];
      }
      __syncthreads();
      for (int t = 0; t < rpt; ++t) {
        const int r = tid + t * kApWG;
        if (r < lf.len) ap_dot<F, NR>(Di + r, a.lm, xs, lf.len, y[t]);
      }
    }
    for (int t = 0; t < rpt; ++t) {
      const int r = tid + t * kApWG;
      if (r < lf.len)
#pragma unroll
        for (int q = 0; q < NR; ++q) a.X[lf.start + r + (int64_t)q * a.ldx] = y[t][q];
    }
    __syncthreads();
  }

  // 2. corrections, deepest parent depth first
  for (int j = L - 1; j >= 0; --j) {
    const ApLevel<F> lv = a.lv[j];
    gptr<F> Q = (gptr<F>)lv.Q;
    // 2a. partials of the owned leaves; the last arrival at a parent solves it
    for (int leaf = blockIdx.x; leaf < a.nleaf; leaf += gridDim.x) {
      const ApEntry en = a.ent[(int64_t)leaf * L + j];
      if (en.p < 0) continue;
      ap_vtx<F, NR>((gptr<F>)lv.V, a.ldv, a.leaves[leaf], en.vcols, a.X, a.ldx, rpt, KMP, red, us);
      ap_publish<F, NR, REFINE>(a, lv, en, leaf, us, ts, cs, us, red, red2, &sh);
    }
    // 2b. wait for the parents' coefficients, X -= Q c over the owned leaves
    for (int leaf = blockIdx.x; leaf < a.nleaf; leaf += gridDim.x) {
      const ApEntry en = a.ent[(int64_t)leaf * L + j];
      if (en.p < 0) continue;
      if (!ap_wait(a.flag, en.p, a.err, a.spin_limit, &sh)) return;
      F* cp = a.coef + (int64_t)en.p * a.MM * NR + (int64_t)en.coff * NR;
      for (int e = tid; e < en.qcols * NR; e += kApWG) cs[e] = ld_dev(cp + e);
      __syncthreads();
      const ApLeaf lf = a.leaves[leaf];
      for (int t = 0; t < rpt; ++t) {
        const int r = tid + t * kApWG;
        if (r >= lf.len) break;
        const int64_t row = lf.start + r;
        F d[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) d[q] = F(0);
        ap_dot<F, NR>(Q + row, a.ldv, cs, en.qcols, d);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          F* xp = a.X + row + (int64_t)q * a.ldx;
          *xp = *xp - d[q];
        }
      }
      __syncthreads();
    }
  }
}

__global__ void k_zero_ints(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

template <class F>
static size_t ap_lds(const ApPlan<F>& p, int nr) {
  const int KMP = (p.KM + 15) / 16 * 16;
  const size_t nred = (size_t)std::max(kApWaves * KMP * nr, kApWG);
  return ((size_t)2 * p.lm * nr + nred + (size_t)(3 + kApWaves) * p.MM * nr) * sizeof(F);
}

template <class F, int NR, bool REFINE>
static int ap_grid_for(const ApPlan<F>& p) {
  int dev = 0, ncu = 0, coop = 0, nb = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) return 0;
  const size_t lds = ap_lds(p, NR);
  if (lds > 64 * 1024) return 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_apply_inverse<F, NR, REFINE>, kApWG, lds));
  return std::min(p.nleaf, nb * ncu);
}

template <class F>
int apply_fused_grid(const ApPlan<F>& p, bool refine, int nr) {
  if (!p.ok || p.lm > (int64_t)kApWG * kApMaxRpt) return 0;
  if (refine) {
    switch (nr) {
      case 1: return ap_grid_for<F, 1, true>(p);
      case 2: return ap_grid_for<F, 2, true>(p);
      default: return ap_grid_for<F, 4, true>(p);
    }
  }
  switch (nr) {
    case 1: return ap_grid_for<F, 1, false>(p);
    case 2: return ap_grid_for<F, 2, false>(p);
    default: return ap_grid_for<F, 4, false>(p);
  }
}

template <class F, int NR, bool REFINE>
static hipError_t ap_launch(ApArgs<F> a, int grid, size_t lds, hipStream_t s) {
  void* args[] = {&a};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_apply_inverse<F, NR, REFINE>), dim3(grid),
                                    dim3(kApWG), args, (unsigned)lds, s);
}

template <class F>
bool apply_inverse_fused(const ApPlan<F>& p, const F* Dinv, const F* D, const F* B, int64_t ldb, int ncols, F* X,
                         int64_t ldx, int64_t ldv, hipStream_t s) {
  const bool refine = D != nullptr;
  for (int nr : {4, 2, 1})
    if (apply_fused_grid(p, refine, nr) <= 0) return false;
  int* sync = p.sync.template as<int>();
  const int nsync = 2 * p.npar + 1;
  int done = 0;
  while (done < ncols) {
    const int rem = ncols - done;
    const int nr = rem >= 4 ? 4 : (rem >= 2 ? 2 : 1);
    const int grid = apply_fused_grid(p, refine, nr);
    ApArgs<F> a;
    a.leaves = p.leaves.get();
    a.ent = p.ent.get();
    a.lv = p.lv.get();
    a.par = p.par.get();
    a.Dinv = Dinv;
    a.D = D;
    a.lm = p.lm;
    a.B = B + (int64_t)done * ldb;
    a.ldb = ldb;
    a.X = X + (int64_t)done * ldx;
    a.ldx = ldx;
    a.ldv = ldv;
    a.part = p.part.template as<F>();
    a.coef = p.coef.template as<F>();
    a.cnt = sync;
    a.flag = sync + p.npar;
    a.err = sync + 2 * p.npar;
    a.nleaf = p.nleaf;
    a.L = p.L;
    a.KM = p.KM;
    a.MM = p.MM;
    a.spin_limit = g_spin_limit.load();
    // counters and flags start at zero (a kernel, not a memset: graph-safe);
    // err is kept, so a failed group is reported after the call
    const int nz = done == 0 ? nsync : 2 * p.npar;  // err is cleared once per call
    k_zero_ints<<<(unsigned)std::max(1, std::min(64, (nz + 255) / 256)), 256, 0, s>>>(sync, nz);
    HIP_CHECK(hipGetLastError());
    const size_t lds = ap_lds(p, nr);
    hipError_t e;
    if (refine) {
      if (nr == 4) e = ap_launch<F, 4, true>(a, grid, lds, s);
      else if (nr == 2) e = ap_launch<F, 2, true>(a, grid, lds, s);
      else e = ap_launch<F, 1, true>(a, grid, lds, s);
    } else {
      if (nr == 4) e = ap_launch<F, 4, false>(a, grid, lds, s);
      else if (nr == 2) e = ap_launch<F, 2, false>(a, grid, lds, s);
      else e = ap_launch<F, 1, false>(a, grid, lds, s);
    }
    if (e != hipSuccess) {
      // the runtime refused the cooperative launch: nothing of this group
      // ran, the caller takes the multi-launch path for every column
      (void)hipGetLastError();
      prof_add("fused_solve_refused", 0.0, 1.0);
      return false;
    }
    done += nr;
  }
  return true;
}

template int apply_fused_grid<double>(const ApPlan<double>&, bool, int);
template int apply_fused_grid<float>(const ApPlan<float>&, bool, int);
template bool apply_inverse_fused<double>(const ApPlan<double>&, const double*, const double*, const double*, int64_t,
                                          int, double*, int64_t, int64_t, hipStream_t);
template bool apply_inverse_fused<float>(const ApPlan<float>&, const float*, const float*, const float*, int64_t, int,
                                         float*, int64_t, int64_t, hipStream_t);

}  // namespace kern
}  // namespace hmat
