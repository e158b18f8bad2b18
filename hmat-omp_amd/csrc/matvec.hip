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
//   reduce   one workgroup per segment: t_s = sum of its leaves' partials;
//   expand   one workgroup per leaf: y = D_leaf x_leaf + sum_c (U S)_c t_sib,
//            x and the t vectors staged in LDS, D and every (U S)_c read
//            coalesced once.
// HBM traffic is the algorithmic 8 (N leaf + 2 N r L) + 16 N nrhs bytes
// plus the O(leaves x depths x r) partials.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"
#include "matvec.hpp"

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
// KG = 16 / NR columns of V: every lane loads its row of the group
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
    // One row per thread.  The steps (depth c, column group kb) run as a
    // pipeline: the next step's loads are issued before this step's
    // reduction, so every wave keeps loads in flight across its shuffles.
    const int r = threadIdx.x;
    const bool live = r < lf.len;
    const int64_t row = lf.start + min(r, lf.len - 1);
    double xr[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) xr[q] = live ? Xg[row + (int64_t)q * ldx] : 0.0;
    auto load = [&](double (&a)[KG], int c, int vcols, int kb) {
      gdouble* V = (gdouble*)lvls[c].V;
      const int64_t N = lvls[c].ld;
#pragma unroll
      for (int kk = 0; kk < KG; ++kk) a[kk] = V[row + (int64_t)min(kb + kk, vcols - 1) * N];
    };
    int c = 0;
    while (c < L && le[c].vcols <= 0) ++c;
    if (c == L) return;
    int vc = le[c].vcols, kb = 0;
    double a[KG];
    load(a, c, vc, kb);
    while (true) {
      int nc = c, nkb = kb + KG, nvc = vc;
      if (nkb >= vc) {
        nkb = 0;
        ++nc;
        while (nc < L && le[nc].vcols <= 0) ++nc;
        if (nc < L) nvc = le[nc].vcols;
      }
      double an[KG];
      if (nc < L) load(an, nc, nvc, nkb);
      double v[16];
#pragma unroll
      for (int kk = 0; kk < KG; ++kk) {
        const double ak = (live && kb + kk < vc) ? a[kk] : 0.0;
#pragma unroll
        for (int q = 0; q < NR; ++q) v[kk * NR + q] = ak * xr[q];
      }
      const double sum = wave_reduce_scatter16(v);  // lane & 15 = kk * NR + q
      if (lane < 16) red[w * KMP * NR + kb * NR + lane] = sum;
      if (nkb == 0) flush(c, vc);
      if (nc >= L) break;
      c = nc;
      kb = nkb;
      vc = nvc;
      for (int kk = 0; kk < KG; ++kk) a[kk] = an[kk];
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
  for (int i0 = 0; i0 < P; i0 += 64) {
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
  for (int r = threadIdx.x; r < lf.len; r += kMvWG) {
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;
    // acc += M(row, j0 .. j0+n) v(j0 .. j0+n): batches of 16 and 8
    // independent loads, then single ones; M column j at M + j * ldm
    auto dot = [&](gdouble* M, int64_t ldm, const double* v, int n) {
      int j = 0;
      for (; j + 16 <= n; j += 16) {
        double m[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) m[u] = M[(int64_t)(j + u) * ldm];
#pragma unroll
        for (int u = 0; u < 16; ++u)
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = fma(m[u], v[(j + u) * NR + q], acc[q]);
      }
      if (j + 8 <= n) {
        double m[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) m[u] = M[(int64_t)(j + u) * ldm];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = fma(m[u], v[(j + u) * NR + q], acc[q]);
        j += 8;
      }
      for (; j < n; ++j) {
        const double m = M[(int64_t)j * ldm];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = fma(m, v[j * NR + q], acc[q]);
      }
    };
    dot(Dl + r, lm, xs, lf.len);
    const int64_t row = lf.start + r;
    for (int c = 0; c < L; ++c) {
      const MvEntry en = ent[(int64_t)blockIdx.x * L + c];
      dot((gdouble*)lvls[c].US + row, lvls[c].ld, ts + (int64_t)c * KM * NR, en.ucols);
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) Y[row + (int64_t)q * ldy] = acc[q];
  }
}

template <int NR>
static void launch_mv(const MvPlan& p, const double* Dm, int64_t lm, const double* X, int64_t ldx,
                      double* part, double* t, double* Y, int64_t ldy, hipStream_t s) {
  const unsigned nl = (unsigned)p.nleaf;
  if (p.L > 0) {
    const size_t plds = (size_t)(kMvWG / 64) * ((p.KM + 15) / 16 * 16) * NR * sizeof(double);
    k_mv_project<NR><<<nl, kMvWG, plds, s>>>(p.leaves, p.lvls, p.ent, p.L, p.KM, X, ldx, part);
    HIP_CHECK(hipGetLastError());
    k_mv_reduce<NR><<<(unsigned)p.nseg, kMvWG, 0, s>>>(p.segs, p.L, p.KM, part, t);
    HIP_CHECK(hipGetLastError());
  }
  const size_t lds = (size_t)(lm + (int64_t)p.L * p.KM) * NR * sizeof(double);
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
    default: throw Error(HMAT_ERR_ARG, "fused matvec takes 1, 2 or 4 right-hand sides");
  }
}

}  // namespace kern
}  // namespace hmat
