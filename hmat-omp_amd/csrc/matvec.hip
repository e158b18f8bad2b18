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

// Lane l ends with the sum over the wave of v[l] (64 values per lane in,
// the lane's own total out): six halving steps of 32 + 16 + ... + 1
// shuffles.  Template recursion keeps every index static, so v stays in
// registers.
template <int S>
__device__ __forceinline__ void halving_step(double (&v)[64], int lane) {
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
__device__ inline double wave_reduce_scatter64(double (&v)[64]) {
  halving_step<32>(v, threadIdx.x & 63);
  return v[0];
}

// partial[((leaf * L + c) * KM + k) * nrhs + col]
template <int NR>
__global__ __launch_bounds__(kMvWG) void k_mv_project(const MvLeaf* __restrict__ leaves,
                                                      const MvLevel* __restrict__ lvls,
                                                      const MvEntry* __restrict__ ent, int L, int KM,
                                                      const double* __restrict__ X, int64_t ldx,
                                                      double* __restrict__ part) {
  constexpr int KG = 64 / NR;  // columns of V per reduction group
  __shared__ double red[kMvWG / 64][64];
  const MvLeaf lf = leaves[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int c = 0; c < L; ++c) {
    const MvEntry e = ent[(int64_t)blockIdx.x * L + c];
    if (e.vcols <= 0) continue;
    const double* V = lvls[c].V;
    const int64_t N = lvls[c].ld;
    for (int kb = 0; kb < e.vcols; kb += KG) {
      double v[64];
#pragma unroll
      for (int i = 0; i < 64; ++i) v[i] = 0.0;
      for (int r = threadIdx.x; r < lf.len; r += kMvWG) {
        const int64_t row = lf.start + r;
        double xr[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) xr[q] = X[row + (int64_t)q * ldx];
#pragma unroll
        for (int kk = 0; kk < KG; ++kk) {
          const int k = kb + kk;
          const double a = k < e.vcols ? V[row + (int64_t)k * N] : 0.0;
#pragma unroll
          for (int q = 0; q < NR; ++q) v[kk * NR + q] = fma(a, xr[q], v[kk * NR + q]);
        }
      }
      const double s = wave_reduce_scatter64(v);  // lane = kk * NR + q
      red[w][lane] = s;
      __syncthreads();
      if (threadIdx.x < 64) {
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < kMvWG / 64; ++i) t += red[i][threadIdx.x];
        const int kk = threadIdx.x / NR, q = threadIdx.x % NR, k = kb + kk;
        if (k < e.vcols) part[(((int64_t)blockIdx.x * L + c) * KM + k) * NR + q] = t;
      }
      __syncthreads();
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
  for (int i0 = 0; i0 < P; i0 += 64) {
    const int i = i0 + lane;
    double acc = 0.0;
    if (i < P)
      for (int leaf = sg.leaf0 + w; leaf < sg.leaf1; leaf += kMvWG / 64)
        acc += part[((int64_t)leaf * L + sg.level) * KM * NR + i];
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
  const double* Dl = Dm + (int64_t)blockIdx.x * lm * lm;
  for (int r = threadIdx.x; r < lf.len; r += kMvWG) {
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0.0;
    int j = 0;
    for (; j + 8 <= lf.len; j += 8) {
      double d[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) d[u] = Dl[r + (int64_t)(j + u) * lm];
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = fma(d[u], xs[(j + u) * NR + q], acc[q]);
    }
    for (; j < lf.len; ++j) {
      const double d = Dl[r + (int64_t)j * lm];
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = fma(d, xs[j * NR + q], acc[q]);
    }
    const int64_t row = lf.start + r;
    for (int c = 0; c < L; ++c) {
      const MvEntry en = ent[(int64_t)blockIdx.x * L + c];
      const double* U = lvls[c].US;
      const int64_t N = lvls[c].ld;
      const double* tc = ts + (int64_t)c * KM * NR;
      int k = 0;
      for (; k + 8 <= en.ucols; k += 8) {
        double u[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) u[i] = U[row + (int64_t)(k + i) * N];
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = fma(u[i], tc[(k + i) * NR + q], acc[q]);
      }
      for (; k < en.ucols; ++k) {
        const double u = U[row + (int64_t)k * N];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = fma(u, tc[k * NR + q], acc[q]);
      }
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
    k_mv_project<NR><<<nl, kMvWG, 0, s>>>(p.leaves, p.lvls, p.ent, p.L, p.KM, X, ldx, part);
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
