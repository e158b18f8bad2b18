// Fused HODLR matvec: y = D x + sum over depths of (U S)_s (V_sib^T x_sib).
//
// Reference: HODLRMatrixOMP::multiply / mult_recursive
// (include/hodlr_routines_omp.h:81-135), which walks the tree with tiny
// GEMVs per node.  Here the whole product is three launches, organised by
// leaf (every leaf lies inside exactly one segment of every depth):
//
//   project  one workgroup per leaf: for each depth c, the leaf's share of
//            V_c^T x over its rows (lanes run down the rows, so V is read
//            coalesced once), reduced across lanes by recursive halving
//            (63 shuffles for 64 values) -> one partial per (leaf, depth);
//            passes of 4, 8 or 16 right-hand sides run it on FP64 MFMA;
//   reduce   one workgroup per 64 values of a segment: t_s = sum of its
//            leaves' partials;
//   expand   one workgroup per leaf: y = D_leaf x_leaf + sum_c (U S)_c t_sib,
//            x and the t vectors staged in LDS, D and every (U S)_c read
//            coalesced once (on FP64 MFMA from four right-hand sides).
// HBM traffic is the algorithmic 8 (N leaf + 2 N r L) + 16 N nrhs bytes
// plus the O(leaves x depths x r) partials.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "matvec.hpp"
#include "mfma.hpp"

namespace hmat {
namespace kern {

constexpr int kMvWG = 256;

// Global-address-space views: loads through them are global_load (not
// flat_load, which also counts against the LDS wait counter).
typedef const __attribute__((address_space(1))) double gdouble;

// 16 values per lane in; lane l ends with the wave total of v[l & 15]:
// four halving steps inside each group of 16 lanes (8 + 4 + 2 + 1
// shuffles), then the four groups are added (2 shuffles).  Template
// recursion keeps the indices static, so v stays in registers.
template <int S>
__device__ __forceinline__ void halving_step(double (&v)[16], int lane) {
  if constexpr (S >= 1) {
    const bool upper = (lane & S) != 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const double send = upper ? v[i] : v[i + S];
      const double keep = upper ? v[i + S] : v[i];
      v[i] = keep + __shfl_xor(send, S);
    }
    halving_step<S / 2>(v, lane);
  }
}
__device__ inline double wave_reduce_scatter16(double (&v)[16]) {
  const int lane = threadIdx.x & 63;
  halving_step<8>(v, lane);
  double s = v[0];
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s;
}

// partial[((leaf * L + c) * KM + k) * nrhs + col].  Per depth, groups of
// KG = 16 / NR columns of V (one column for NR = 16): every lane loads its row of the group
// (unconditionally, the column index clamped), the wave reduces, and the
// four waves meet in LDS once per depth.
template <int NR>
__global__ __launch_bounds__(kMvWG) void k_mv_project(const MvLeaf* __restrict__ leaves,
                                                      const MvLevel* __restrict__ lvls,
                                                      const MvEntry* __restrict__ ent, int L, int KM,
                                                      const double* __restrict__ X, int64_t ldx,
                                                      double* __restrict__ part) {
  constexpr int KG = 16 / NR;
  // [4][KMP * NR]: KMP = KM rounded up to 16, so the last group's 16 lanes
  // stay inside their wave's section
  extern __shared__ double red[];
  const int KMP = (KM + 15) / 16 * 16;
  const MvLeaf lf = leaves[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  gdouble* Xg = (gdouble*)X;
  const MvEntry* le = ent + (int64_t)blockIdx.x * L;
  auto flush = [&](int c, int vcols) {  // the four waves' sums of depth c -> partials
    __syncthreads();
    for (int i = threadIdx.x; i < vcols * NR; i += kMvWG) {
      double t = 0.0;
#pragma unroll
      for (int j = 0; j < kMvWG / 64; ++j) t += red[j * KMP * NR + i];
      part[((int64_t)blockIdx.x * L + c) * KM * NR + i] = t;
    }
    __syncthreads();
  };
  if (lf.len <= kMvWG) {
    // One row per thread.  A step loads 16 columns of V (the same depth of
    // loads for every NR) and reduces them in 16 / KG rounds of KG columns
    // x NR right-hand sides.  The steps (depth c, column block kb) run as a
    // pipeline: the next step's loads are issued before this step's
    // reductions, so every wave keeps loads in flight across its shuffles.
    constexpr int LU = 16;
    const int r = threadIdx.x;
    const bool live = r < lf.len;
    const int64_t row = lf.start + min(r, lf.len - 1);
    double xr[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) xr[q] = live ? Xg[row + (int64_t)q * ldx] : 0.0;
    auto load = [&](double (&a)[LU], int c, int vcols, int kb) {
      gdouble* V = (gdouble*)lvls[c].V;
      const int64_t N = lvls[c].ld;
#pragma unroll
      for (int kk = 0; kk < LU; ++kk) a[kk] = V[row + (int64_t)min(kb + kk, vcols - 1) * N];
    };
    int c = 0;
    while (c < L && le[c].vcols <= 0) ++c;
    if (c == L) return;
    int vc = le[c].vcols, kb = 0;
    double a[LU];
    load(a, c, vc, kb);
    while (true) {
      int nc = c, nkb = kb + LU, nvc = vc;
      if (nkb >= vc) {
        nkb = 0;
        ++nc;
        while (nc < L && le[nc].vcols <= 0) ++nc;
        if (nc < L) nvc = le[nc].vcols;
      }
      double an[LU];
      if (nc < L) load(an, nc, nvc, nkb);
#pragma unroll
      for (int g = 0; g < LU / KG; ++g) {
        if (kb + g * KG < vc) {  // the same for the whole workgroup
          double v[16];
#pragma unroll
          for (int kk = 0; kk < KG; ++kk) {
            const double ak = (live && kb + g * KG + kk < vc) ? a[g * KG + kk] : 0.0;
#pragma unroll
            for (int q = 0; q < NR; ++q) v[kk * NR + q] = ak * xr[q];
          }
          const double sum = wave_reduce_scatter16(v);  // lane & 15 = kk * NR + q
          if (lane < 16) red[w * KMP * NR + (kb + g * KG) * NR + lane] = sum;
        }
      }
      if (nkb == 0) flush(c, vc);
      if (nc >= L) break;
      c = nc;
      kb = nkb;
      vc = nvc;
      for (int kk = 0; kk < LU; ++kk) a[kk] = an[kk];
    }
    return;
  }
  for (int c = 0; c < L; ++c) {
    const MvEntry e = ent[(int64_t)blockIdx.x * L + c];
    if (e.vcols <= 0) continue;
    gdouble* V = (gdouble*)lvls[c].V;
    const int64_t N = lvls[c].ld;
    for (int kb = 0; kb < e.vcols; kb += KG) {
      double v[16];
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = 0.0;
      for (int r = threadIdx.x; r < lf.len; r += kMvWG) {
        const int64_t row = lf.start + r;
        double xr[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) xr[q] = Xg[row + (int64_t)q * ldx];
        double a[KG];
#pragma unroll
        for (int kk = 0; kk < KG; ++kk) a[kk] = V[row + (int64_t)min(kb + kk, e.vcols - 1) * N];
#pragma unroll
        for (int kk = 0; kk < KG; ++kk) {
          const double ak = kb + kk < e.vcols ? a[kk] : 0.0;
#pragma unroll
          for (int q = 0; q < NR; ++q) v[kk * NR + q] = fma(ak, xr[q], v[kk * NR + q]);
        }
      }
      const double s = wave_reduce_scatter16(v);  // lane & 15 = kk * NR + q
      if (lane < 16) red[w * KMP * NR + kb * NR + lane] = s;
    }
    flush(c, e.vcols);
  }
}

// Orders this wave's LDS writes before its later LDS reads of other
// lanes' data (and the reads before later writes): a wave's LDS operations
// run in issue order, so a compiler barrier suffices.
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MFMA passes (leaves of at most 256 rows; see launch_mv): the per-leaf
// projection is a small GEMM, (V_c^T)(rows of the leaf) X(rows, 0:NR), run
// on v_mfma_f64_16x16x4f64.  The leaf's X rows are staged in LDS as 16
// columns (zero beyond NR and past the leaf's rows); the leaf's (depth,
// 16-column block) items are dealt round-robin to the waves of gridDim.y
// workgroups.  A wave reads its item's 16 V columns 64 rows at a time with
// the lanes down the rows (coalesced, sixteen loads per lane in flight
// during the previous chunk's products), transposes them through its own
// LDS tile into the MFMA operand layout (lane l takes column l % 16, rows
// l / 16 + 4u), and writes the 16 x NR block of partials.  Same partial
// layout as k_mv_project.
constexpr int kMvVS = 66;  // LDS stride of a staged V column (doubles): the operand reads of one
                           // half-wave (columns 0..15, two rows) touch 64 distinct banks
constexpr size_t kMvProjTile = (size_t)(kMvWG / 64) * 16 * kMvVS;  // doubles
template <int NR>
__global__ __launch_bounds__(kMvWG) void k_mv_project_mfma(const MvLeaf* __restrict__ leaves,
                                                           const MvLevel* __restrict__ lvls,
                                                           const MvEntry* __restrict__ ent, int L, int KM,
                                                           const double* __restrict__ X, int64_t ldx,
                                                           double* __restrict__ part) {
    extern __shared__ double sm[];  // per wave 16 x kMvVS, then x: lenp x 16
  const MvLeaf lf = leaves[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double* vsw = sm + (size_t)w * 16 * kMvVS;
  double* xs = sm + kMvProjTile;
  const int lenp = (lf.len + 3) & ~3;
  for (int e = threadIdx.x; e < lenp * 16; e += kMvWG) {
    const int r = e / 16, q = e % 16;
    xs[e] = (r < lf.len && q < NR) ? X[lf.start + r + (int64_t)q * ldx] : 0.0;
  }
  __syncthreads();
  const MvEntry* le = ent + (int64_t)blockIdx.x * L;
  const int kr = lane >> 4, ci = lane & 15;
  const int waves = (kMvWG / 64) * (int)gridDim.y, mine = (int)blockIdx.y * (kMvWG / 64) + w;
  const int nch = (lf.len + 63) / 64;
  int item = 0;
  for (int c = 0; c < L; ++c) {
    const int vc = le[c].vcols;
    if (vc <= 0) continue;
    gdouble* V = (gdouble*)lvls[c].V;
    const int64_t N = lvls[c].ld;
    for (int cb = 0; cb < vc; cb += 16, ++item) {
      if (item % waves != mine) continue;
      double g[16];
      auto load = [&](int ch) {  // rows and columns clamped: every load in bounds
        const int64_t row = lf.start + min(ch * 64 + lane, lf.len - 1);
#pragma unroll
        for (int j = 0; j < 16; ++j) g[j] = V[row + (int64_t)min(cb + j, vc - 1) * N];
      };
      load(0);
      mfma::v4d acc = {0.0, 0.0, 0.0, 0.0};
      for (int ch = 0; ch < nch; ++ch) {
        const bool rok = ch * 64 + lane < lf.len;
        wave_lds_sync();  // the previous chunk's operand reads are done
#pragma unroll
        for (int j = 0; j < 16; ++j) vsw[j * kMvVS + lane] = (rok && cb + j < vc) ? g[j] : 0.0;
        wave_lds_sync();
        if (ch + 1 < nch) load(ch + 1);
        const int r0 = ch * 64;
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (r0 + 4 * u < lenp) {  // rows up to lenp: x is staged (zero past the leaf)
            const int r = 4 * u + kr;
            acc = mfma::f64_16x16x4(vsw[ci * kMvVS + r], xs[(r0 + r) * 16 + ci], acc);
          }
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = mfma::f64_row(lane, v), q = mfma::f64_col(lane, v);
        if (cb + i < vc && q < NR) part[(((int64_t)blockIdx.x * L + c) * KM + cb + i) * NR + q] = acc[v];
      }
    }
  }
}

// t[(g * KM + k) * nrhs + col] = sum over the segment's leaves of the partials
template <int NR>
__global__ __launch_bounds__(kMvWG) void k_mv_reduce(const MvSeg* __restrict__ segs, int L, int KM,
                                                     const double* __restrict__ part,
                                                     double* __restrict__ t) {
  __shared__ double red[kMvWG / 64][64];
  const MvSeg sg = segs[blockIdx.x];
  const int P = sg.vcols * NR;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int WS = kMvWG / 64;  // wave w sums leaves w, w + WS, w + 2 WS, ...
  const int64_t lstride = (int64_t)L * KM * NR;
  const int nlf = sg.leaf1 - sg.leaf0;
  // gridDim.y workgroups share the segment's values, 64 at a time
  for (int i0 = (int)blockIdx.y * 64; i0 < P; i0 += 64 * (int)gridDim.y) {
    const int i = i0 + lane;
    double acc = 0.0;
    if (i < P) {
      const double* pp = part + ((int64_t)sg.leaf0 * L + sg.level) * KM * NR + i;
      int l = w;
      // eight independent loads in flight; the additions keep their order
      for (; l + 7 * WS < nlf; l += 8 * WS) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = pp[(int64_t)(l + u * WS) * lstride];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += v[u];
      }
      for (; l < nlf; l += WS) acc += pp[(int64_t)l * lstride];
    }
    red[w][lane] = acc;
    __syncthreads();
    if (threadIdx.x < 64 && i < P) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < kMvWG / 64; ++j) s += red[j][threadIdx.x];
      t[(int64_t)blockIdx.x * KM * NR + i] = s;
    }
    __syncthreads();
  }
}

// y(leaf rows) = D_leaf x(leaf rows) + sum_c (U S)_c(leaf rows, :) t_{c, sib}
template <int NR>
__global__ __launch_bounds__(kMvWG) void k_mv_expand(const MvLeaf* __restrict__ leaves,
                                                     const MvLevel* __restrict__ lvls,
                                                     const MvEntry* __restrict__ ent, int L, int KM,
                                                     const double* __restrict__ Dm, int64_t lm,
                                                     const double* __restrict__ X, int64_t ldx,
                                                     const double* __restrict__ t,
                                                     double* __restrict__ Y, int64_t ldy) {
  extern __shared__ double sm[];  // x: lm x NR, then per depth KM x NR
  const MvLeaf lf = leaves[blockIdx.x];
  double* xs = sm;
  double* ts = sm + lm * NR;
  for (int e = threadIdx.x; e < lf.len * NR; e += kMvWG) {
    const int r = e / NR, q = e % NR;
    xs[e] = X[lf.start + r + (int64_t)q * ldx];
  }
  for (int c = 0; c < L; ++c) {
    const MvEntry en = ent[(int64_t)blockIdx.x * L + c];
    for (int e = threadIdx.x; e < en.ucols * NR; e += kMvWG)
      ts[(int64_t)c * KM * NR + e] = t[(int64_t)en.tin * KM * NR + e];
  }
  __syncthreads();
  gdouble* Dl = (gdouble*)(Dm + (int64_t)blockIdx.x * lm * lm);
  // Short leaves leave threads without a row: there each row's products
  // are split over tpr adjacent lanes (contiguous column ranges), summed by
  // shuffles, so every wave streams loads.
  const int tpr = lf.len <= 64 ? 4 : (lf.len <= 128 ? 2 : 1);
  const int sub = threadIdx.x % tpr;
  const int rpp = kMvWG / tpr;  // rows per pass
  for (int r0 = 0; r0 < lf.len; r0 += rpp) {
    const int r = r0 + threadIdx.x / tpr;
    const bool live = r < lf.len;
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;
    // acc += M(row, j0 .. j0+n) v(j0 .. j0+n): batches of 16 and 8
    // independent loads (8 or 4 for the wide passes, which keep NR
    // accumulators), then single ones; M column j at M + j * ldm
    auto dot = [&](gdouble* M, int64_t ldm, const double* v, int n) {
      int j = 0;
      if constexpr (NR <= 4)
        for (; j + 16 <= n; j += 16) {
          double m[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) m[u] = M[(int64_t)(j + u) * ldm];
#pragma unroll
          for (int u = 0; u < 16; ++u)
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] = fma(m[u], v[(j + u) * NR + q], acc[q]);
        }
      constexpr int B = NR <= 8 ? 8 : 4;
      for (; j + B <= n; j += B) {
        double m[B];
#pragma unroll
        for (int u = 0; u < B; ++u) m[u] = M[(int64_t)(j + u) * ldm];
#pragma unroll
        for (int u = 0; u < B; ++u)
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = fma(m[u], v[(j + u) * NR + q], acc[q]);
      }
      for (; j < n; ++j) {
        const double m = M[(int64_t)j * ldm];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = fma(m, v[j * NR + q], acc[q]);
      }
    };
    // this lane's share [j0, j1) of n columns
    auto part_dot = [&](gdouble* M, int64_t ldm, const double* v, int n) {
      const int j0 = (int)((int64_t)n * sub / tpr), j1 = (int)((int64_t)n * (sub + 1) / tpr);
      dot(M + (int64_t)j0 * ldm, ldm, v + (int64_t)j0 * NR, j1 - j0);
    };
    const int64_t row = lf.start + (live ? r : 0);
    if (live) {
      part_dot(Dl + r, lm, xs, lf.len);
      for (int c = 0; c < L; ++c) {
        const MvEntry en = ent[(int64_t)blockIdx.x * L + c];
        part_dot((gdouble*)lvls[c].US + row, lvls[c].ld, ts + (int64_t)c * KM * NR, en.ucols);
      }
    }
    for (int o = 1; o < tpr; o <<= 1)
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] += __shfl_xor(acc[q], o);
    if (live && sub == 0)
#pragma unroll
      for (int q = 0; q < NR; ++q) Y[row + (int64_t)q * ldy] = acc[q];
  }
}

// MFMA passes (leaves of at most 256 rows; see launch_mv): the same product
// as k_mv_expand, y(leaf rows, 0:NR) = [D_leaf | (U S)_c ...] [x; t_sib ...],
// as a small GEMM on v_mfma_f64_16x16x4f64 with the same LDS staging.  Each
// wave takes two 16-row tiles at a time and streams their rows of every
// matrix four columns per MFMA, sixteen loads per lane in flight (the VALU
// form keeps four per lane from four right-hand sides).
template <int NR>
__global__ __launch_bounds__(kMvWG) void k_mv_expand_mfma(const MvLeaf* __restrict__ leaves,
                                                          const MvLevel* __restrict__ lvls,
                                                          const MvEntry* __restrict__ ent, int L, int KM,
                                                          const double* __restrict__ Dm, int64_t lm,
                                                          const double* __restrict__ X, int64_t ldx,
                                                          const double* __restrict__ t,
                                                          double* __restrict__ Y, int64_t ldy) {
    extern __shared__ double sm[];  // x: lm x NR, then per depth KM x NR
  const MvLeaf lf = leaves[blockIdx.x];
  double* xs = sm;
  double* ts = sm + lm * NR;
  for (int e = threadIdx.x; e < lf.len * NR; e += kMvWG) {
    const int r = e / NR, q = e % NR;
    xs[e] = X[lf.start + r + (int64_t)q * ldx];
  }
  for (int c = 0; c < L; ++c) {
    const MvEntry en = ent[(int64_t)blockIdx.x * L + c];
    for (int e = threadIdx.x; e < en.ucols * NR; e += kMvWG)
      ts[(int64_t)c * KM * NR + e] = t[(int64_t)en.tin * KM * NR + e];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, kr = lane >> 4;
  const int qb = min(i, NR - 1);  // B column of this lane (zero beyond NR)
  const int ntile = (lf.len + 15) / 16;
  gdouble* Dl = (gdouble*)(Dm + (int64_t)blockIdx.x * lm * lm);
  for (int t0 = w; t0 < ntile; t0 += 2 * (kMvWG / 64)) {
    const int t1 = t0 + kMvWG / 64;  // past the last tile: computed on clamped rows, not stored
    const int ra = min(t0 * 16 + i, lf.len - 1), rb = min(t1 * 16 + i, lf.len - 1);
    mfma::v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    // acc += M(rows, 0:n) v(0:n, 0:NR); M column j at M + j * ldm, rows
    // relative to the leaf start
    auto seg = [&](gdouble* M, int64_t ldm, const double* v, int n) {
      for (int k0 = 0; k0 < n; k0 += 32) {
        double a0[8], a1[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int k = k0 + 4 * u + kr;
          const int kc = min(k, n - 1);
          const double m0 = M[ra + (int64_t)kc * ldm];
          const double m1 = M[rb + (int64_t)kc * ldm];
          a0[u] = k < n ? m0 : 0.0;
          a1[u] = k < n ? m1 : 0.0;
          b[u] = (k < n && i < NR) ? v[kc * NR + qb] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (k0 + 4 * u < n) {
            acc0 = mfma::f64_16x16x4(a0[u], b[u], acc0);
            acc1 = mfma::f64_16x16x4(a1[u], b[u], acc1);
          }
      }
    };
    seg(Dl, lm, xs, lf.len);
    for (int c = 0; c < L; ++c) {
      const MvEntry en = ent[(int64_t)blockIdx.x * L + c];
      if (en.ucols > 0) seg((gdouble*)lvls[c].US + lf.start, lvls[c].ld, ts + (int64_t)c * KM * NR, en.ucols);
    }
    if (i < NR) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = mfma::f64_row(lane, v);
        const int r0 = t0 * 16 + r, r1 = t1 * 16 + r;
        if (r0 < lf.len) Y[lf.start + r0 + (int64_t)i * ldy] = acc0[v];
        if (t1 < ntile && r1 < lf.len) Y[lf.start + r1 + (int64_t)i * ldy] = acc1[v];
      }
    }
  }
}

template <int NR>
static void launch_mv(const MvPlan& p, const double* Dm, int64_t lm, const double* X, int64_t ldx,
                      double* part, double* t, double* Y, int64_t ldy, hipStream_t s) {
  const unsigned nl = (unsigned)p.nleaf;
  // MFMA passes from four columns, and at every width on leaves of at most
  // 128 rows, where the row-per-thread projection leaves half of its
  // workgroup idle (one column, 65536 rows in leaves of 128: 0.097 ms
  // against 0.108 ms; in leaves of 256 the VALU kernels stay ahead for one
  // column, 0.471 ms against 0.492 ms at 2^18 rows)
  const bool mf = lm <= kMvWG && (NR >= 4 || lm <= 128);
  if (p.L > 0) {
    if (mf) {
      const size_t xl = (kMvProjTile + (size_t)((lm + 3) & ~3) * 16) * sizeof(double);
      // up to four workgroups per leaf, so that a few hundred leaves still
      // put about 1024 workgroups in flight
      const unsigned sy = (unsigned)std::max(1, std::min(4, (1024 + p.nleaf - 1) / std::max(p.nleaf, 1)));
      k_mv_project_mfma<NR><<<dim3(nl, sy), kMvWG, xl, s>>>(p.leaves, p.lvls, p.ent, p.L, p.KM, X, ldx, part);
    } else {
      const size_t plds = (size_t)(kMvWG / 64) * ((p.KM + 15) / 16 * 16) * NR * sizeof(double);
      k_mv_project<NR><<<nl, kMvWG, plds, s>>>(p.leaves, p.lvls, p.ent, p.L, p.KM, X, ldx, part);
    }
    HIP_CHECK(hipGetLastError());
    // one workgroup per 64 values of a segment's t
    const unsigned vy = (unsigned)std::max(1, (p.KM * NR + 63) / 64);
    k_mv_reduce<NR><<<dim3((unsigned)p.nseg, vy), kMvWG, 0, s>>>(p.segs, p.L, p.KM, part, t);
    HIP_CHECK(hipGetLastError());
  }
  const size_t lds = (size_t)(lm + (int64_t)p.L * p.KM) * NR * sizeof(double);
  if (mf)
    k_mv_expand_mfma<NR><<<nl, kMvWG, lds, s>>>(p.leaves, p.lvls, p.ent, p.L, p.KM, Dm, lm, X, ldx, t, Y, ldy);
  else
    k_mv_expand<NR><<<nl, kMvWG, lds, s>>>(p.leaves, p.lvls, p.ent, p.L, p.KM, Dm, lm, X, ldx, t, Y, ldy);
  HIP_CHECK(hipGetLastError());
}

size_t mv_lds_bytes(const MvPlan& p, int64_t lm, int nrhs) {
  return (size_t)(lm + (int64_t)p.L * p.KM) * nrhs * sizeof(double);
}

void hodlr_matvec_fused(const MvPlan& p, const double* Dm, int64_t lm, const double* X, int64_t ldx,
                        int nrhs, double* part, double* t, double* Y, int64_t ldy, hipStream_t s) {
  switch (nrhs) {
    case 1: launch_mv<1>(p, Dm, lm, X, ldx, part, t, Y, ldy, s); break;
    case 2: launch_mv<2>(p, Dm, lm, X, ldx, part, t, Y, ldy, s); break;
    case 4: launch_mv<4>(p, Dm, lm, X, ldx, part, t, Y, ldy, s); break;
    case 8: launch_mv<8>(p, Dm, lm, X, ldx, part, t, Y, ldy, s); break;
    case 16: launch_mv<16>(p, Dm, lm, X, ldx, part, t, Y, ldy, s); break;
    default: throw Error(HMAT_ERR_ARG, "fused matvec takes 1, 2, 4, 8 or 16 right-hand sides");
  }
}

}  // namespace kern
}  // namespace hmat
