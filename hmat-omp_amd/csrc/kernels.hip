// gfx950 kernels of the HODLR / HSS engine (FP64 path).
//
// Wavefront = 64 lanes; workgroups are 256 threads (4 waves).  The dense
// contraction of the build (the randomised sketch of every off-diagonal
// block of a tree level) runs on FP64 MFMA (v_mfma_f64_16x16x4f64) with the
// kernel-matrix entries generated in registers, so A never exists in HBM.
// Everything else is tall-skinny and bandwidth-bound.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "generators.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace hmat {
namespace kern {

constexpr int kWG = 256;

// --------------------------------------------------------------------------
// seg_gemm_nn: Y(seg) = alpha X(seg) T_s + beta Y(seg)
// One workgroup owns R rows x (G*16) columns of one segment.  Lanes run
// down the rows (coalesced X and Y); the small per-segment T is read with
// wave-uniform addresses (scalar loads).
// --------------------------------------------------------------------------

template <class F, int R>
__global__ __launch_bounds__(kWG) void k_seg_gemm_nn(const SegNN* __restrict__ segs,
                                                     const NNTile* __restrict__ tiles,
                                                     const F* __restrict__ X, int64_t ldx,
                                                     const F* __restrict__ T, int64_t ldt,
                                                     F* __restrict__ Y, int64_t ldy,
                                                     F alpha, F beta) {
  constexpr int G = kWG / R;
  constexpr int CPT = 16;
  const NNTile tl = tiles[blockIdx.x];
  const SegNN sg = segs[tl.seg];
  const int r = threadIdx.x % R;
  const int g = __builtin_amdgcn_readfirstlane(threadIdx.x / R);
  const int row = tl.row0 + r;
  const bool live = row < sg.len;
  const int ncols = min(G * CPT, sg.q - tl.c0);
  F acc[CPT];
#pragma unroll
  for (int i = 0; i < CPT; ++i) acc[i] = F(0);
  const F* xr = X + sg.xoff + (live ? row : 0);
  const F* tb = T + sg.toff + (int64_t)(tl.c0 + g) * ldt;
  // eight independent X loads in flight per step (a one-load loop waits on
  // every load); the accumulation order over i is unchanged
  int i = 0;
  for (; i + 8 <= sg.p; i += 8) {
    F xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = xr[(int64_t)(i + u) * ldx];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (!live) xv[u] = F(0);
#pragma unroll
      for (int cc = 0; cc < CPT; ++cc)
        if (g + cc * G < ncols) acc[cc] = fma(xv[u], tb[i + u + (int64_t)(cc * G) * ldt], acc[cc]);
    }
  }
  for (; i < sg.p; ++i) {
    const F xv = live ? xr[(int64_t)i * ldx] : F(0);
#pragma unroll
    for (int cc = 0; cc < CPT; ++cc) {
      if (g + cc * G < ncols) acc[cc] = fma(xv, tb[i + (int64_t)(cc * G) * ldt], acc[cc]);
    }
  }
  if (!live) return;
#pragma unroll
  for (int cc = 0; cc < CPT; ++cc) {
    const int c = g + cc * G;
    if (c < ncols) {
      F* yp = Y + sg.yoff + row + (int64_t)(tl.c0 + c) * ldy;
      F v = alpha * acc[cc];
      if (beta != F(0)) v = fma(beta, *yp, v);
      *yp = v;
    }
  }
}

// MFMA variant for wide segments (p >= 16 and q >= 16, e.g. the leaf
// blocks D_i (len x len) applied to many columns): 64 x 64 output tile per
// workgroup, each wave a 32 x 32 quarter as 2 x 2 MFMA 16x16 tiles; K-tiles
// of 16 staged in LDS.  FP64: v_mfma_f64_16x16x4f64, FP32:
// v_mfma_f32_16x16x4f32.
template <class F>
struct MfmaNN;
template <>
struct MfmaNN<double> {
  typedef mfma::v4d acc_t;
  static __device__ acc_t step(double a, double b, acc_t c) { return mfma::f64_16x16x4(a, b, c); }
  static __device__ int row(int lane, int v) { return mfma::f64_row(lane, v); }
};
template <>
struct MfmaNN<float> {
  typedef mfma::v4f acc_t;
  static __device__ acc_t step(float a, float b, acc_t c) { return mfma::f32_16x16x4(a, b, c); }
  static __device__ int row(int lane, int v) { return mfma::f32_row(lane, v); }
};

template <class F>
__global__ __launch_bounds__(kWG) void k_seg_gemm_nn_mfma(const SegNN* __restrict__ segs,
                                                          const NNTile* __restrict__ tiles,
                                                          const F* __restrict__ X, int64_t ldx,
                                                          const F* __restrict__ T, int64_t ldt,
                                                          F* __restrict__ Y, int64_t ldy, F alpha,
                                                          F beta) {
  constexpr int KT = 16, TM = 64, TN = 64;
  __shared__ F xs[KT][TM + 1];  // xs[k][r] = X(row0 + r, k0 + k)
  __shared__ F ts[KT][TN + 1];  // ts[k][c] = T(k0 + k, c0 + c)
  const NNTile tl = tiles[blockIdx.x];
  const SegNN sg = segs[tl.seg];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int wm = w & 1, wn = w >> 1;
  typedef typename MfmaNN<F>::acc_t acc_t;
  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = acc_t{F(0), F(0), F(0), F(0)};
  for (int k0 = 0; k0 < sg.p; k0 += KT) {
    for (int e = threadIdx.x; e < KT * TM; e += kWG) {
      const int r = e % TM, k = e / TM;
      xs[k][r] = (tl.row0 + r < sg.len && k0 + k < sg.p) ? X[sg.xoff + tl.row0 + r + (int64_t)(k0 + k) * ldx]
                                                         : F(0);
    }
    for (int e = threadIdx.x; e < KT * TN; e += kWG) {
      const int k = e % KT, c = e / KT;
      ts[k][c] = (k0 + k < sg.p && tl.c0 + c < sg.q) ? T[sg.toff + k0 + k + (int64_t)(tl.c0 + c) * ldt] : F(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KT; kk += 4) {
      F a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = xs[kk + lane / 16][wm * 32 + i * 16 + lane % 16];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = ts[kk + lane / 16][wn * 32 + j * 16 + lane % 16];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = MfmaNN<F>::step(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = tl.row0 + wm * 32 + i * 16 + MfmaNN<F>::row(lane, v);
        const int c = tl.c0 + wn * 32 + j * 16 + (lane & 15);
        if (r < sg.len && c < sg.q) {
          F* yp = Y + sg.yoff + r + (int64_t)c * ldy;
          F val = alpha * acc[i][j][v];
          if (beta != F(0)) val = fma(beta, *yp, val);
          *yp = val;
        }
      }
}

// Tiling of a segmented Y = alpha X T + beta Y: the kernel variant (1: MFMA
// 64 x 64 tiles, else rows per workgroup) and its tiles.
int nn_tiles(const std::vector<SegNN>& segs, std::vector<NNTile>& tiles) {
  tiles.clear();
  int qmax = 0;
  for (auto& g : segs) qmax = std::max(qmax, g.q);
  if (qmax == 0) return 0;
  int pmin = 1 << 30;
  for (auto& g : segs)
    if (g.len > 0 && g.q > 0) pmin = std::min(pmin, g.p);
  int mode, R, cw;
  if (pmin >= 16 && qmax >= 16) {
    mode = 1;  // MFMA, 64 x 64 tiles (K-tiles of 16, the last zero-padded)
    R = 64;
    cw = 64;
  } else {
    R = qmax <= 16 ? 256 : (qmax <= 32 ? 128 : 64);
    mode = R;
    cw = (kWG / R) * 16;
  }
  for (size_t i = 0; i < segs.size(); ++i) {
    const auto& g = segs[i];
    if (g.len <= 0 || g.q <= 0) continue;
    for (int r0 = 0; r0 < g.len; r0 += R)
      for (int c0 = 0; c0 < g.q; c0 += cw) tiles.push_back({(int32_t)i, r0, c0, 0});
  }
  return mode;
}

template <class F>
void launch_nn_raw(int mode, int ntiles, const SegNN* ds, const NNTile* dt, const F* X, int64_t ldx,
                          const F* T, int64_t ldt, F* Y, int64_t ldy, F alpha, F beta, hipStream_t s) {
  if (!ntiles) return;
  const unsigned g = (unsigned)ntiles;
  if (mode == 1)
    k_seg_gemm_nn_mfma<F><<<g, kWG, 0, s>>>(ds, dt, X, ldx, T, ldt, Y, ldy, alpha, beta);
  else if (mode == 256)
    k_seg_gemm_nn<F, 256><<<g, kWG, 0, s>>>(ds, dt, X, ldx, T, ldt, Y, ldy, alpha, beta);
  else if (mode == 128)
    k_seg_gemm_nn<F, 128><<<g, kWG, 0, s>>>(ds, dt, X, ldx, T, ldt, Y, ldy, alpha, beta);
  else
    k_seg_gemm_nn<F, 64><<<g, kWG, 0, s>>>(ds, dt, X, ldx, T, ldt, Y, ldy, alpha, beta);
  HIP_CHECK(hipGetLastError());
}

void plan_nn(const std::vector<SegNN>& segs, NNPlan& p, hipStream_t s) {
  std::vector<NNTile> tiles;
  p.mode = segs.empty() ? 0 : nn_tiles(segs, tiles);
  p.ntiles = (int)tiles.size();
  if (!p.ntiles) return;
  p.segs.upload(segs, s);
  p.tiles.upload(tiles, s);
}

template <class F>
void launch_nn(const NNPlan& p, const F* X, int64_t ldx, const F* T, int64_t ldt, F* Y, int64_t ldy, F alpha,
               F beta, hipStream_t s) {
  launch_nn_raw<F>(p.mode, p.ntiles, p.segs.get(), p.tiles.get(), X, ldx, T, ldt, Y, ldy, alpha, beta, s);
}
template void launch_nn_raw<double>(int, int, const SegNN*, const NNTile*, const double*, int64_t, const double*,
                                    int64_t, double*, int64_t, double, double, hipStream_t);
template void launch_nn_raw<float>(int, int, const SegNN*, const NNTile*, const float*, int64_t, const float*,
                                   int64_t, float*, int64_t, float, float, hipStream_t);
template void launch_nn<double>(const NNPlan&, const double*, int64_t, const double*, int64_t, double*, int64_t,
                                double, double, hipStream_t);
template void launch_nn<float>(const NNPlan&, const float*, int64_t, const float*, int64_t, float*, int64_t, float,
                               float, hipStream_t);

template <class F>
void seg_gemm_nn(const std::vector<SegNN>& segs, const F* X, int64_t ldx,
                 const F* T, int64_t ldt, F* Y, int64_t ldy, F alpha, F beta, hipStream_t s) {
  if (segs.empty()) return;
  std::vector<NNTile> tiles;
  const int mode = nn_tiles(segs, tiles);
  if (tiles.empty()) return;
  StageGroup g;
  const int is = g.add(segs), it = g.add(tiles);
  g.put(s);
  launch_nn_raw<F>(mode, (int)tiles.size(), g.get<SegNN>(is), g.get<NNTile>(it), X, ldx, T, ldt, Y, ldy, alpha,
                   beta, s);
}
template void seg_gemm_nn<double>(const std::vector<SegNN>&, const double*, int64_t,
                                  const double*, int64_t, double*, int64_t, double, double, hipStream_t);
template void seg_gemm_nn<float>(const std::vector<SegNN>&, const float*, int64_t,
                                 const float*, int64_t, float*, int64_t, float, float, hipStream_t);

// --------------------------------------------------------------------------
// seg_trsm_upper: Y(seg) = X(seg) R_s^{-1}, R_s upper triangular (zero
// pivots give zero columns).  Forward substitution along each row.
//
// k <= 64: the row lives in registers (K = 16/32/64 doubles per lane, padded
// with R = I beyond k), R in LDS and read with wave-uniform (broadcast)
// addresses; 256 rows per workgroup, so X is read and Y written exactly once
// with coalesced column accesses.  k > 64: the row lives in LDS.
// --------------------------------------------------------------------------
// Straight-line substitution, generated by template recursion so that the
// K(K-1)/2 updates are fully unrolled for every K (the loop unroller gives
// up at K = 64, which would put the row in scratch memory).
template <int K, int J, int C0>
__device__ __forceinline__ void trsm_chunk(double (&q)[K], const double* rt) {
  if constexpr (C0 < K) {
    constexpr int CE = (C0 + 8 < K) ? C0 + 8 : K;
    double r[8];
#pragma unroll
    for (int c = C0; c < CE; ++c) r[c - C0] = rt[J * K + c];  // the chunk's loads go out together
#pragma unroll
    for (int c = C0; c < CE; ++c) q[c] = fma(-q[J], r[c - C0], q[c]);
    // the chunk's results are final before the next chunk's R loads are
    // issued, so at most 8 R values are live (otherwise every load is
    // hoisted up front)
#pragma unroll
    for (int c = C0; c < CE; ++c) asm volatile("" : "+v"(q[c]));
    asm volatile("" ::: "memory");
    trsm_chunk<K, J, C0 + 8>(q, rt);
  }
}
template <int K, int J>
__device__ __forceinline__ void trsm_step(double (&q)[K], const double* rt) {
  if constexpr (J < K) {
    const double d = rt[J * K + J];
    q[J] = d != 0.0 ? q[J] / d : 0.0;
    trsm_chunk<K, J, J + 1>(q, rt);
    trsm_step<K, J + 1>(q, rt);
  }
}

template <int K>
__global__ __launch_bounds__(kWG) void k_seg_trsm_reg(const SegNN* __restrict__ segs,
                                                      const NNTile* __restrict__ tiles,
                                                      const double* __restrict__ X, int64_t ldx,
                                                      const double* __restrict__ R, int64_t ldr,
                                                      double* __restrict__ Y, int64_t ldy) {
  __shared__ double rt[K * K];  // R^T: row j of R at rt[j*K]
  const NNTile tl = tiles[blockIdx.x];
  const SegNN sg = segs[tl.seg];
  const int k = sg.q;
  for (int e = threadIdx.x; e < K * K; e += kWG) {
    const int c = e % K, j = e / K;  // rt[j*K + c] = R(j, c)
    rt[e] = (j < k && c < k) ? R[sg.toff + j + (int64_t)c * ldr] : (j == c ? 1.0 : 0.0);
  }
  const int row = tl.row0 + threadIdx.x;
  const bool live = row < sg.len;
  double q[K];
#pragma unroll
  for (int j = 0; j < K; ++j) q[j] = (live && j < k) ? X[sg.xoff + row + (int64_t)j * ldx] : 0.0;
  __syncthreads();
  // right-looking: once q_j is final, q_c -= q_j R(j, c) for c > j.  Each
  // q_c sees the same operations in the same order as the left-looking
  // recurrence q_c = (x_c - sum_{i<c} q_i R(i, c)) / R(c, c).
  trsm_step<K, 0>(q, rt);
  if (!live) return;
#pragma unroll
  for (int j = 0; j < K; ++j)
    if (j < k) Y[sg.yoff + row + (int64_t)j * ldy] = q[j];
}

__global__ __launch_bounds__(64) void k_seg_trsm_upper(const SegNN* __restrict__ segs,
                                                       const NNTile* __restrict__ tiles,
                                                       const double* __restrict__ X, int64_t ldx,
                                                       const double* __restrict__ R, int64_t ldr,
                                                       double* __restrict__ Y, int64_t ldy,
                                                       int r_in_lds) {
  extern __shared__ double q[];  // [k][64], then R (k x k) when it fits
  const NNTile tl = tiles[blockIdx.x];
  const SegNN sg = segs[tl.seg];
  const int t = threadIdx.x;
  const int row = tl.row0 + t;
  const bool live = row < sg.len;
  const int k = sg.q;
  for (int j = 0; j < k; ++j) q[j * 64 + t] = live ? X[sg.xoff + row + (int64_t)j * ldx] : 0.0;
  const double* rbase = R + sg.toff;
  int64_t lrr = ldr;
  if (r_in_lds) {
    double* rs = q + (int64_t)k * 64;
    for (int e = t; e < k * k; e += 64) rs[e] = R[sg.toff + e % k + (int64_t)(e / k) * ldr];
    __syncthreads();
    rbase = rs;
    lrr = k;
  }
  for (int j = 0; j < k; ++j) {
    const double* rj = rbase + (int64_t)j * lrr;
    double acc = q[j * 64 + t];
    for (int i = 0; i < j; ++i) acc = fma(-q[i * 64 + t], rj[i], acc);
    q[j * 64 + t] = rj[j] != 0.0 ? acc / rj[j] : 0.0;
  }
  if (!live) return;
  for (int j = 0; j < k; ++j) Y[sg.yoff + row + (int64_t)j * ldy] = q[j * 64 + t];
}

int trsm_tiles(const std::vector<SegNN>& segs, std::vector<NNTile>& tiles) {
  tiles.clear();
  int kmax = 0;
  for (const auto& g : segs)
    if (g.len > 0 && g.q > 0) kmax = std::max(kmax, g.q);
  if (kmax == 0) return 0;
  HMAT_REQUIRE(kmax <= 300, HMAT_ERR_ARG, "seg_trsm_upper: k > 300");
  const int rows = kmax <= 64 ? kWG : 64;
  for (size_t i = 0; i < segs.size(); ++i) {
    const auto& g = segs[i];
    if (g.len <= 0 || g.q <= 0) continue;
    for (int r0 = 0; r0 < g.len; r0 += rows) tiles.push_back({(int32_t)i, r0, 0, 0});
  }
  return kmax;
}

void launch_trsm_raw(int kmax, int ntiles, const SegNN* ds, const NNTile* dt, const double* X, int64_t ldx,
                            const double* R, int64_t ldr, double* Y, int64_t ldy, hipStream_t s) {
  if (!ntiles) return;
  const unsigned nt = (unsigned)ntiles;
  if (kmax <= 16) {
    k_seg_trsm_reg<16><<<nt, kWG, 0, s>>>(ds, dt, X, ldx, R, ldr, Y, ldy);
  } else if (kmax <= 32) {
    k_seg_trsm_reg<32><<<nt, kWG, 0, s>>>(ds, dt, X, ldx, R, ldr, Y, ldy);
  } else if (kmax <= 48) {
    k_seg_trsm_reg<48><<<nt, kWG, 0, s>>>(ds, dt, X, ldx, R, ldr, Y, ldy);
  } else if (kmax <= 64) {
    k_seg_trsm_reg<64><<<nt, kWG, 0, s>>>(ds, dt, X, ldx, R, ldr, Y, ldy);
  } else {
    size_t lds = (size_t)kmax * 64 * sizeof(double);
    const size_t rlds = (size_t)kmax * kmax * sizeof(double);
    const int r_in_lds = (lds + rlds <= 96 * 1024) ? 1 : 0;
    if (r_in_lds) lds += rlds;
    k_seg_trsm_upper<<<nt, 64, lds, s>>>(ds, dt, X, ldx, R, ldr, Y, ldy, r_in_lds);
  }
  HIP_CHECK(hipGetLastError());
}

void seg_trsm_upper(const std::vector<SegNN>& segs, const double* X,
                    int64_t ldx, const double* R, int64_t ldr, double* Y, int64_t ldy, hipStream_t s) {
  if (segs.empty()) return;
  std::vector<NNTile> tiles;
  const int kmax = trsm_tiles(segs, tiles);
  if (!kmax) return;
  StageGroup g;
  const int is = g.add(segs), it = g.add(tiles);
  g.put(s);
  launch_trsm_raw(kmax, (int)tiles.size(), g.get<SegNN>(is), g.get<NNTile>(it), X, ldx, R, ldr, Y, ldy, s);
}

void plan_trsm(const std::vector<SegNN>& segs, TrsmPlan& p, hipStream_t s) {
  std::vector<NNTile> tiles;
  p.kmax = segs.empty() ? 0 : trsm_tiles(segs, tiles);
  p.ntiles = (int)tiles.size();
  if (!p.ntiles) return;
  p.segs.upload(segs, s);
  p.tiles.upload(tiles, s);
}

void launch_trsm(const TrsmPlan& p, const double* X, int64_t ldx, const double* R, int64_t ldr, double* Y,
                 int64_t ldy, hipStream_t s) {
  launch_trsm_raw(p.kmax, p.ntiles, p.segs.get(), p.tiles.get(), X, ldx, R, ldr, Y, ldy, s);
}

// --------------------------------------------------------------------------
// seg_gemm_tn: O_s = alpha X(seg)^T Y(seg) + beta O_s, two stages:
// partial 64x64 products over row chunks, then a deterministic sum.
// --------------------------------------------------------------------------

template <class F>
__global__ __launch_bounds__(kWG) void k_seg_gemm_tn_reduce(const SegTN* __restrict__ segs,
                                                            const TNOut* __restrict__ outs,
                                                            const F* __restrict__ part,
                                                            F* __restrict__ O, int64_t ldo,
                                                            F alpha, F beta) {
  // blockIdx.y picks 64 of the 4096 entries of the output tile; the four
  // waves sum interleaved quarters of the slots, then wave 0 adds the
  // quarters in a fixed order (deterministic)
  __shared__ F red[4][64];
  const TNOut ot = outs[blockIdx.x];
  const SegTN sg = segs[ot.seg];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int e = blockIdx.y * 64 + lane;
  F sum = F(0);
  for (int c = w; c < ot.nslot; c += 4) sum += part[(int64_t)(ot.slot0 + c) * 4096 + e];
  red[w][lane] = sum;
  __syncthreads();
  if (w != 0) return;
  const int i = e % 64, j = e / 64;
  const int gi = ot.pt * 64 + i, gj = ot.qt * 64 + j;
  if (gi >= sg.p || gj >= sg.q) return;
  F* o = O + sg.ooff + gi + (int64_t)gj * ldo;
  F v = alpha * ((red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]));
  if (beta != F(0)) v = fma(beta, *o, v);
  *o = v;
}

// MFMA partial products: one 64 x 64 output tile over a row chunk, staged
// through LDS 32 rows at a time with the next 32 rows prefetched into
// registers; wave w computes output rows 16w..16w+15 (columns of X) against
// all four 16-column groups of Y with 16x16x4 MFMAs.  When X and Y are the
// same rows and columns (a Gram matrix) the chunk is read once.  An output
// tile covered by one chunk (slot < 0) is written to O directly, exactly as
// the sum stage would write it; the others go to their partial slot.
template <class F>
__global__ __launch_bounds__(kWG) void k_seg_gemm_tn_part_mfma(const SegTN* __restrict__ segs,
                                                               const TNItem* __restrict__ items,
                                                               const F* __restrict__ X, int64_t ldx,
                                                               const F* __restrict__ Y, int64_t ldy,
                                                               F* __restrict__ part, F* __restrict__ O,
                                                               int64_t ldo, F alpha, F beta) {
  constexpr int KR = 32, NPRE = KR * 64 / kWG;
  __shared__ F xs[KR][65];
  __shared__ F ys[KR][65];
  const TNItem it = items[blockIdx.x];
  const SegTN sg = segs[it.seg];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const bool same = X == Y && ldx == ldy && sg.xoff == sg.yoff && sg.p == sg.q && it.pt == it.qt;
  typedef typename MfmaNN<F>::acc_t acc_t;
  acc_t acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = acc_t{F(0), F(0), F(0), F(0)};
  const int pc0 = it.pt * 64, qc0 = it.qt * 64;
  F px[NPRE], py[NPRE];
  auto issue = [&](int rb) {
#pragma unroll
    for (int i = 0; i < NPRE; ++i) {
      const int e = threadIdx.x + i * kWG;
      const int rr = e % KR, cc = e / KR, row = rb + rr;
      const bool rok = row < it.r1;
      px[i] = (rok && pc0 + cc < sg.p) ? X[sg.xoff + row + (int64_t)(pc0 + cc) * ldx] : F(0);
      if (!same) py[i] = (rok && qc0 + cc < sg.q) ? Y[sg.yoff + row + (int64_t)(qc0 + cc) * ldy] : F(0);
    }
  };
  auto commit = [&]() {
#pragma unroll
    for (int i = 0; i < NPRE; ++i) {
      const int e = threadIdx.x + i * kWG;
      xs[e % KR][e / KR] = px[i];
      if (!same) ys[e % KR][e / KR] = py[i];
    }
  };
  F(*bs)[65] = same ? xs : ys;
  issue(it.r0);
  commit();
  __syncthreads();
  for (int rb = it.r0; rb < it.r1; rb += KR) {
    const bool more = rb + KR < it.r1;
    if (more) issue(rb + KR);
#pragma unroll
    for (int kk = 0; kk < KR; kk += 4) {
      const F a = xs[kk + lane / 16][16 * w + lane % 16];
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = MfmaNN<F>::step(a, bs[kk + lane / 16][16 * t + lane % 16], acc[t]);
    }
    __syncthreads();
    if (more) {
      commit();
      __syncthreads();
    }
  }
  if (it.slot < 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gi = pc0 + 16 * w + MfmaNN<F>::row(lane, v), gj = qc0 + 16 * t + (lane & 15);
        if (gi < sg.p && gj < sg.q) {
          F* o = O + sg.ooff + gi + (int64_t)gj * ldo;
          F val = alpha * acc[t][v];
          if (beta != F(0)) val = fma(beta, *o, val);
          *o = val;
        }
      }
    return;
  }
  F* pp = part + (int64_t)it.slot * 4096;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) pp[(16 * w + MfmaNN<F>::row(lane, v)) + (16 * t + (lane & 15)) * 64] = acc[t][v];
}

// Work items of a segmented O = alpha X^T Y + beta O: partial 64 x 64
// products over row chunks (slots of the partial buffer) and their sums.
// An output tile of one chunk needs no sum: its item writes O (slot -1).
int tn_items(const std::vector<SegTN>& segs, std::vector<TNItem>& items, std::vector<TNOut>& outs) {
  items.clear();
  outs.clear();
  int64_t total_rows = 0;
  for (auto& g : segs) total_rows += std::max(g.len, 0);
  // ~1024 partial items in total, chunks a multiple of 32 rows
  int64_t ch = std::max<int64_t>(256, round_up((int)std::min<int64_t>(ceil_div(total_rows, 1024), 1 << 20), 32));
  int slot = 0;
  for (size_t i = 0; i < segs.size(); ++i) {
    const auto& g = segs[i];
    if (g.p <= 0 || g.q <= 0) continue;
    const int npt = (int)ceil_div(g.p, 64), nqt = (int)ceil_div(g.q, 64);
    for (int pt = 0; pt < npt; ++pt)
      for (int qt = 0; qt < nqt; ++qt) {
        if (g.len <= ch) {
          items.push_back({(int32_t)i, 0, (int32_t)std::max<int64_t>(g.len, 0), pt, qt, -1});
          continue;
        }
        TNOut o{(int32_t)i, pt, qt, slot, 0, 0};
        for (int64_t r0 = 0; r0 < g.len; r0 += ch) {
          items.push_back({(int32_t)i, (int32_t)r0, (int32_t)std::min<int64_t>(r0 + ch, g.len), pt, qt, slot++});
          o.nslot++;
        }
        outs.push_back(o);
      }
  }
  return slot;
}

template <class F>
void launch_tn_raw(int nitems, int nouts, const SegTN* ds, const TNItem* di, const TNOut* dout,
                          const F* X, int64_t ldx, const F* Y, int64_t ldy, F* O, int64_t ldo, F alpha, F beta,
                          F* part, hipStream_t s) {
  if (!nitems) return;
  k_seg_gemm_tn_part_mfma<F><<<(unsigned)nitems, kWG, 0, s>>>(ds, di, X, ldx, Y, ldy, part, O, ldo, alpha, beta);
  HIP_CHECK(hipGetLastError());
  if (!nouts) return;
  k_seg_gemm_tn_reduce<F><<<dim3((unsigned)nouts, 4096 / 64), kWG, 0, s>>>(ds, dout, part, O, ldo, alpha, beta);
  HIP_CHECK(hipGetLastError());
}

void plan_tn(const std::vector<SegTN>& segs, TNPlan& p, hipStream_t s) {
  p.nitems = p.nouts = p.nslots = 0;
  if (segs.empty()) return;
  std::vector<TNItem> items;
  std::vector<TNOut> outs;
  const int nslots = tn_items(segs, items, outs);
  if (items.empty()) return;
  p.nitems = (int)items.size();
  p.nouts = (int)outs.size();
  p.nslots = nslots;
  p.segs.upload(segs, s);
  p.items.upload(items, s);
  p.outs.upload(outs, s);
}

template <class F>
void launch_tn(const TNPlan& p, const F* X, int64_t ldx, const F* Y, int64_t ldy, F* O, int64_t ldo, F alpha,
               F beta, F* part, hipStream_t s) {
  launch_tn_raw<F>(p.nitems, p.nouts, p.segs.get(), p.items.get(), p.outs.get(), X, ldx, Y, ldy, O, ldo, alpha,
                   beta, part, s);
}
template void launch_tn_raw<double>(int, int, const SegTN*, const TNItem*, const TNOut*, const double*, int64_t,
                                    const double*, int64_t, double*, int64_t, double, double, double*, hipStream_t);
template void launch_tn_raw<float>(int, int, const SegTN*, const TNItem*, const TNOut*, const float*, int64_t,
                                   const float*, int64_t, float*, int64_t, float, float, float*, hipStream_t);
template void launch_tn<double>(const TNPlan&, const double*, int64_t, const double*, int64_t, double*, int64_t,
                                double, double, double*, hipStream_t);
template void launch_tn<float>(const TNPlan&, const float*, int64_t, const float*, int64_t, float*, int64_t,
                               float, float, float*, hipStream_t);

template <class F>
void seg_gemm_tn(const std::vector<SegTN>& segs, const F* X, int64_t ldx,
                 const F* Y, int64_t ldy, F* O, int64_t ldo, F alpha, F beta, Workspace& ws,
                 hipStream_t s) {
  if (segs.empty()) return;
  std::vector<TNItem> items;
  std::vector<TNOut> outs;
  const int nslots = tn_items(segs, items, outs);
  if (items.empty()) return;
  DevBuf<F>& part = ws.parts<F>();
  part.reserve((size_t)std::max(nslots, 1) * 4096);
  StageGroup g;
  const int is = g.add(segs), ii = g.add(items), io = g.add(outs);
  g.put(s);
  launch_tn_raw<F>((int)items.size(), (int)outs.size(), g.get<SegTN>(is), g.get<TNItem>(ii), g.get<TNOut>(io), X,
                   ldx, Y, ldy, O, ldo, alpha, beta, part.get(), s);
}
template void seg_gemm_tn<double>(const std::vector<SegTN>&, const double*, int64_t,
                                  const double*, int64_t, double*, int64_t, double, double,
                                  Workspace&, hipStream_t);
template void seg_gemm_tn<float>(const std::vector<SegTN>&, const float*, int64_t,
                                 const float*, int64_t, float*, int64_t, float, float, Workspace&,
                                 hipStream_t);

// --------------------------------------------------------------------------
// Batched small dense kernels: one workgroup per matrix, matrices in global
// memory (they stay in L1/L2 for the sizes used here, n <= a few hundred).
// --------------------------------------------------------------------------
__device__ inline double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kWG / 64; ++i) t += red[i];
  return t;
}

// R^T R = G + shift I (upper R).  A zero Gram (all-zero input) returns
// R = 0, which the triangular solve turns into a zero basis.  The matrix is
// factored in LDS (dynamic, n*(n+1) doubles) when it fits, else in place in
// global memory.
__global__ __launch_bounds__(kWG) void k_chol(const SmallMat* __restrict__ mats,
                                              const double* __restrict__ G, double* __restrict__ R,
                                              const double* __restrict__ shift_coef,
                                              int* __restrict__ info, double* __restrict__ gerr,
                                              int use_lds, const double* __restrict__ fallback_coef,
                                              int* __restrict__ fb_used) {
  extern __shared__ double sm[];
  __shared__ double red[kWG / 64];
  __shared__ double dred[kWG / 64];
  const SmallMat m = mats[blockIdx.x];
  const int n = m.n, ld = m.ld;
  const double* g = G + m.off;
  double* rg = R + m.off;
  double* r = use_lds ? sm : rg;
  const int lr = use_lds ? n + 1 : ld;
  double tr = 0.0, dev = 0.0;
  for (int i = threadIdx.x; i < n; i += kWG) tr += g[i + (int64_t)i * ld];
  // deviation from I over the nonzero columns: a column that is exactly
  // zero (G_jj = 0, hence G_ij = 0 for every i) stays zero through every
  // pass and carries nothing into the projection
  for (int e = threadIdx.x; e < n * n; e += kWG) {
    const int i = e % n, j = e / n;
    if (g[i + (int64_t)i * ld] != 0.0 && g[j + (int64_t)j * ld] != 0.0)
      dev = fmax(dev, fabs(g[i + (int64_t)j * ld] - (i == j ? 1.0 : 0.0)));
  }
  tr = block_sum(tr, red);
  for (int o = 32; o > 0; o >>= 1) dev = fmax(dev, __shfl_xor(dev, o));
  if (threadIdx.x % 64 == 0) dred[threadIdx.x / 64] = dev;
  __syncthreads();
  if (threadIdx.x == 0) {
    double d = 0.0;
    for (int i = 0; i < kWG / 64; ++i) d = fmax(d, dred[i]);
    gerr[blockIdx.x] = (tr == 0.0) ? 0.0 : d;
    info[blockIdx.x] = 0;
    if (fb_used) fb_used[blockIdx.x] = 0;
  }
  if (!(tr > 0.0)) {
    if (!(tr == 0.0) && threadIdx.x == 0) info[blockIdx.x] = -1;  // NaN
    for (int e = threadIdx.x; e < n * n; e += kWG) rg[e % n + (int64_t)(e / n) * ld] = 0.0;
    return;
  }
  double shift = shift_coef[blockIdx.x] * tr;
  // an unshifted factorisation that breaks down is redone with the
  // fallback shift (when one is given), in this launch
  const double fb = (fallback_coef && shift == 0.0) ? fallback_coef[blockIdx.x] * tr : 0.0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    for (int e = threadIdx.x; e < n * n; e += kWG) {
      const int i = e % n, j = e / n;
      r[i + (int64_t)j * lr] = (i <= j) ? g[i + (int64_t)j * ld] + (i == j ? shift : 0.0) : 0.0;
    }
    __syncthreads();
    int bad = 0;
    for (int j = 0; j < n; ++j) {
      const double d = r[j + (int64_t)j * lr];
      if (!(d > 0.0) || !isfinite(d)) {
        bad = j + 1;
        break;
      }
      const double rjj = sqrt(d);
      __syncthreads();
      for (int c = j + 1 + threadIdx.x; c < n; c += kWG) r[j + (int64_t)c * lr] /= rjj;
      if (threadIdx.x == 0) r[j + (int64_t)j * lr] = rjj;
      __syncthreads();
      const int rem = n - j - 1;
      for (int e = threadIdx.x; e < rem * rem; e += kWG) {
        const int a2 = j + 1 + e % rem, c = j + 1 + e / rem;
        if (a2 <= c) r[a2 + (int64_t)c * lr] -= r[j + (int64_t)a2 * lr] * r[j + (int64_t)c * lr];
      }
      __syncthreads();
    }
    __syncthreads();
    if (!bad) break;
    if (attempt == 1 || !(fb > 0.0)) {
      if (threadIdx.x == 0) info[blockIdx.x] = bad;
      break;
    }
    shift = fb;
    if (fb_used && threadIdx.x == 0) fb_used[blockIdx.x] = 1;
  }
  __syncthreads();
  if (use_lds)
    for (int e = threadIdx.x; e < n * n; e += kWG) {
      const int i = e % n, j = e / n;
      rg[i + (int64_t)j * ld] = r[i + (int64_t)j * lr];
    }
}

// n <= 64, registers: one wave per matrix, lane c holds column c of R
// (col[i] = R(i, c)).  Step j: R(j, j) = sqrt of lane j's col[j] (read by
// every lane with readlane), lanes c > j scale their col[j], and each
// col[a] (a > j) takes -R(j, a) R(j, c) with R(j, a) read from lane a.
// All indices are static (template recursion), so the column stays in
// registers and the loop needs no barrier.  Same outputs and breakdown
// rules as k_chol.
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

template <int J, int A>
__device__ __forceinline__ void chol_update(double (&col)[64], double rjc) {
  if constexpr (A < 64) {
    const double rja = readlane_f64(col[J], A);  // R(J, A) from lane A
    col[A] = fma(-rja, rjc, col[A]);
    chol_update<J, A + 1>(col, rjc);
  }
}

template <int J>
__device__ __forceinline__ int chol_step(double (&col)[64], int n, int c) {
  if constexpr (J < 64) {
    if (J >= n) return 0;
    const double d = readlane_f64(col[J], J);
    if (!(d > 0.0) || !isfinite(d)) return J + 1;
    const double rjj = sqrt(d);
    double rjc = 0.0;
    if (c == J) col[J] = rjj;
    if (c > J) {
      col[J] /= rjj;
      rjc = col[J];
    }
    chol_update<J, J + 1>(col, rjc);
    return chol_step<J + 1>(col, n, c);
  }
  return 0;
}

__global__ __launch_bounds__(64) void k_chol64r(const SmallMat* __restrict__ mats,
                                                const double* __restrict__ G, double* __restrict__ R,
                                                const double* __restrict__ shift_coef,
                                                int* __restrict__ info, double* __restrict__ gerr, int lg,
                                                const double* __restrict__ fallback_coef,
                                                int* __restrict__ fb_used) {
  // G staged through LDS with coalesced loads, column j of G at gs[j * lg];
  // lg = the batch's largest n + 1, so small matrices leave room for more
  // workgroups per CU
  extern __shared__ double gs[];
  const SmallMat m = mats[blockIdx.x];
  const int n = m.n, ld = m.ld;
  const double* g = G + m.off;
  double* rg = R + m.off;
  const int c = threadIdx.x;
  for (int j = 0; j < n; ++j)
    if (c < n) gs[j * lg + c] = g[c + (int64_t)j * ld];
  __syncthreads();
  double tr = 0.0, dev = 0.0;
  if (c < n) {
    tr = gs[c * lg + c];
    // nonzero columns only (an exactly zero column stays zero, see k_chol)
    if (tr != 0.0)
      for (int i = 0; i < n; ++i)
        if (gs[i * lg + i] != 0.0) dev = fmax(dev, fabs(gs[c * lg + i] - (i == c ? 1.0 : 0.0)));
  }
  for (int o = 32; o > 0; o >>= 1) {
    tr += __shfl_xor(tr, o);
    dev = fmax(dev, __shfl_xor(dev, o));
  }
  if (c == 0) {
    gerr[blockIdx.x] = (tr == 0.0) ? 0.0 : dev;
    info[blockIdx.x] = 0;
    if (fb_used) fb_used[blockIdx.x] = 0;
  }
  double shift = shift_coef[blockIdx.x] * tr;
  // an unshifted factorisation that breaks down is redone with the
  // fallback shift (when one is given), in this launch
  const double fb = (fallback_coef && shift == 0.0) ? fallback_coef[blockIdx.x] * tr : 0.0;
  double col[64];
  if (!(tr > 0.0)) {
    if (!(tr == 0.0) && c == 0) info[blockIdx.x] = -1;  // NaN
    if (c < n)
      for (int i = 0; i < n; ++i) rg[i + (int64_t)c * ld] = 0.0;
    return;
  }
  int bad = 0;
#pragma nounroll
  for (int attempt = 0; attempt < 2; ++attempt) {
#pragma unroll
    for (int i = 0; i < 64; ++i) col[i] = (c < n && i <= c) ? gs[c * lg + i] + (i == c ? shift : 0.0) : 0.0;
    bad = chol_step<0>(col, n, c);
    if (!bad || !(fb > 0.0)) break;
    shift = fb;
    if (fb_used && c == 0) fb_used[blockIdx.x] = 1;
  }
  if (bad) {
    if (c == 0) info[blockIdx.x] = bad;
    return;
  }
  if (c < n) {
#pragma unroll
    for (int i = 0; i < 64; ++i)
      if (i < n) rg[i + (int64_t)c * ld] = i <= c ? col[i] : 0.0;
  }
}

// n <= 64: one wave per matrix, lane c owns column c of R (in LDS); the
// column loop needs no cross-wave barriers.  Same outputs and breakdown
// rules as k_chol.
__global__ __launch_bounds__(64) void k_chol64(const SmallMat* __restrict__ mats,
                                               const double* __restrict__ G, double* __restrict__ R,
                                               const double* __restrict__ shift_coef,
                                               int* __restrict__ info, double* __restrict__ gerr) {
  extern __shared__ double r[];  // n x (n+1), column c at r[c * (n+1)]
  const SmallMat m = mats[blockIdx.x];
  const int n = m.n, ld = m.ld, lr = n + 1;
  const double* g = G + m.off;
  double* rg = R + m.off;
  const int c = threadIdx.x;
  double tr = 0.0, dev = 0.0;
  if (c < n) {
    tr = g[c + (int64_t)c * ld];
    // nonzero columns only (an exactly zero column stays zero, see k_chol)
    if (tr != 0.0)
      for (int i = 0; i < n; ++i)
        if (g[i + (int64_t)i * ld] != 0.0) dev = fmax(dev, fabs(g[i + (int64_t)c * ld] - (i == c ? 1.0 : 0.0)));
  }
  for (int o = 32; o > 0; o >>= 1) {
    tr += __shfl_xor(tr, o);
    dev = fmax(dev, __shfl_xor(dev, o));
  }
  if (c == 0) {
    gerr[blockIdx.x] = (tr == 0.0) ? 0.0 : dev;
    info[blockIdx.x] = 0;
  }
  const double shift = shift_coef[blockIdx.x] * tr;
  if (c < n)
    for (int i = 0; i < n; ++i) r[i + c * lr] = (i <= c) ? g[i + (int64_t)c * ld] + (i == c ? shift : 0.0) : 0.0;
  if (!(tr > 0.0)) {
    if (!(tr == 0.0) && c == 0) info[blockIdx.x] = -1;  // NaN
    if (c < n)
      for (int i = 0; i < n; ++i) rg[i + (int64_t)c * ld] = 0.0;
    return;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const double d = r[j + j * lr];
    if (!(d > 0.0) || !isfinite(d)) {
      if (c == 0) info[blockIdx.x] = j + 1;
      break;
    }
    const double rjj = sqrt(d);
    double rjc = 0.0;
    if (c > j && c < n) {
      rjc = r[j + c * lr] / rjj;
      r[j + c * lr] = rjc;
    }
    if (c == j) r[j + j * lr] = rjj;
    __syncthreads();
    // column c: r(a, c) -= r(j, a) r(j, c) for j < a <= c
    if (c > j && c < n)
      for (int a2 = j + 1; a2 <= c; ++a2) r[a2 + c * lr] -= r[j + a2 * lr] * rjc;
    __syncthreads();
  }
  if (c < n)
    for (int i = 0; i < n; ++i) rg[i + (int64_t)c * ld] = r[i + c * lr];
}

void launch_chol_shifted(const SmallMat* dmats, int nmats, int nmax, double* G, double* R, const double* dshift,
                         int* info, double* gerr, hipStream_t s, const double* dfallback) {
  if (!nmats) return;
  const size_t lds = (size_t)nmax * (nmax + 1) * sizeof(double);
  const int use_lds = lds <= 64 * 1024 ? 1 : 0;
  // with profiling on, every matrix whose factorisation took the fallback
  // shift is counted ("chol_fallbacks" over "chol_fallback_mats")
  DevBuf<int> fb;
  if (dfallback && prof_enabled()) fb.resize(nmats);
  if (nmax <= 64)
    k_chol64r<<<(unsigned)nmats, 64, (size_t)nmax * (nmax + 1) * sizeof(double), s>>>(
        dmats, G, R, dshift, info, gerr, nmax + 1, dfallback, fb.get());
  else
    k_chol<<<(unsigned)nmats, kWG, use_lds ? lds : 0, s>>>(dmats, G, R, dshift, info, gerr, use_lds, dfallback,
                                                           fb.get());
  HIP_CHECK(hipGetLastError());
  if (fb.get()) prof_defer_count("chol_fallbacks", "chol_fallback_mats", std::move(fb), (size_t)nmats, s);
}

void chol_shifted(const std::vector<SmallMat>& mats, double* G, double* R, const std::vector<double>& shift_coef,
                  int* info, double* gerr, hipStream_t s, const std::vector<double>& fallback_coef) {
  if (mats.empty()) return;
  HMAT_REQUIRE(fallback_coef.empty() || fallback_coef.size() == mats.size(), HMAT_ERR_ARG,
               "chol_shifted: one fallback coefficient per matrix");
  int nmax = 0;
  for (auto& mm : mats) nmax = std::max(nmax, mm.n);
  StageGroup g;
  const int im = g.add(mats), ic = g.add(shift_coef), ifb = g.add(fallback_coef);
  g.put(s);
  launch_chol_shifted(g.get<SmallMat>(im), (int)mats.size(), nmax, G, R, g.get<double>(ic), info, gerr, s,
                      g.get<double>(ifb));
}

// One-sided Jacobi (Hestenes) with a round-robin pair schedule; the four
// waves rotate disjoint column pairs in parallel.  A and V live in LDS
// (dynamic, 2 n (n+1) doubles) when they fit, else in global memory.
__global__ __launch_bounds__(kWG) void k_jacobi_svd(const SmallMat* __restrict__ mats,
                                                    double* __restrict__ A, double* __restrict__ V,
                                                    double* __restrict__ S, int64_t sstride,
                                                    double* __restrict__ work, int use_lds,
                                                    int* __restrict__ sweeps_out, double floor_rel) {
  extern __shared__ double sm[];
  __shared__ int rotated;
  __shared__ double sig[512];
  __shared__ int perm[512];
  const SmallMat m = mats[blockIdx.x];
  const int n = m.n, ld = m.ld;
  double* ag = A + m.off;
  double* vg = V + m.off;
  double* a = use_lds ? sm : ag;
  double* v = use_lds ? sm + (int64_t)n * (n + 1) : vg;
  const int la = use_lds ? n + 1 : ld;
  double* sv = S + (int64_t)blockIdx.x * sstride;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  double fro = 0.0;
  for (int e = threadIdx.x; e < n * n; e += kWG) {
    const int i = e % n, j = e / n;
    const double x = ag[i + (int64_t)j * ld];
    fro += x * x;
    if (use_lds) a[i + (int64_t)j * la] = x;
    v[i + (int64_t)j * la] = (i == j) ? 1.0 : 0.0;
  }
  // ||A||_F^2: pairs whose inner product is below 1e-30 ||A||_F^2 (about
  // 20 eps^2) are rounding noise of the rotations and are left alone, so
  // columns at the noise level do not keep the sweep loop alive.
  fro = block_sum(fro, sig);
  const double floor2 = floor_rel * fro;
  __syncthreads();
  const int np = n + (n & 1);  // even player count; column n is a phantom
  const int circ = np - 1;
  int sweep = 0;
  for (; sweep < 60 && n > 1; ++sweep) {
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int step = 0; step < circ; ++step) {
      for (int pi = w; pi < np / 2; pi += kWG / 64) {
        int p, q;
        if (pi == 0) {
          p = 0;
          q = 1 + step % circ;
        } else {
          p = 1 + (step + pi) % circ;
          q = 1 + (step - pi + circ) % circ;
        }
        if (p >= n || q >= n) continue;
        if (p > q) { const int t = p; p = q; q = t; }
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = lane; i < n; i += 64) {
          const double x = a[i + (int64_t)p * la], y = a[i + (int64_t)q * la];
          al += x * x;
          be += y * y;
          ga += x * y;
        }
        for (int o = 32; o > 0; o >>= 1) {
          al += __shfl_xor(al, o);
          be += __shfl_xor(be, o);
          ga += __shfl_xor(ga, o);
        }
        if (!(fabs(ga) > 1e-15 * sqrt(al * be)) || !(fabs(ga) > floor2)) continue;
        const double zeta = (be - al) / (2.0 * ga);
        double t;
        if (fabs(zeta) > 1e150)
          t = 0.5 / zeta;
        else
          t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
        for (int i = lane; i < n; i += 64) {
          const double x = a[i + (int64_t)p * la], y = a[i + (int64_t)q * la];
          a[i + (int64_t)p * la] = c * x - sn * y;
          a[i + (int64_t)q * la] = sn * x + c * y;
          const double vx = v[i + (int64_t)p * la], vy = v[i + (int64_t)q * la];
          v[i + (int64_t)p * la] = c * vx - sn * vy;
          v[i + (int64_t)q * la] = sn * vx + c * vy;
        }
        if (lane == 0) rotated = 1;
      }
      __syncthreads();
    }
    if (!rotated) break;
    __syncthreads();
  }
  if (sweeps_out && threadIdx.x == 0) sweeps_out[blockIdx.x] = sweep;
  // singular values = column norms
  for (int j = w; j < n; j += kWG / 64) {
    double s2 = 0.0;
    for (int i = lane; i < n; i += 64) s2 += a[i + (int64_t)j * la] * a[i + (int64_t)j * la];
    for (int o = 32; o > 0; o >>= 1) s2 += __shfl_xor(s2, o);
    if (lane == 0) sig[j] = sqrt(s2);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 0; j < n; ++j) perm[j] = j;
    for (int j = 0; j < n; ++j) {  // selection sort, descending
      int best = j;
      for (int t = j + 1; t < n; ++t)
        if (sig[perm[t]] > sig[perm[best]]) best = t;
      const int tmp = perm[j];
      perm[j] = perm[best];
      perm[best] = tmp;
    }
  }
  __syncthreads();
  const double* su = a;
  const double* svv = v;
  int ls = la;
  if (!use_lds) {
    double* wu = work + m.off * 2;
    double* wv = wu + (int64_t)n * ld;
    for (int e = threadIdx.x; e < n * n; e += kWG) {
      const int i = e % n, j = e / n;
      wu[i + (int64_t)j * ld] = a[i + (int64_t)j * ld];
      wv[i + (int64_t)j * ld] = v[i + (int64_t)j * ld];
    }
    __syncthreads();
    su = wu;
    svv = wv;
    ls = ld;
  }
  for (int e = threadIdx.x; e < n * n; e += kWG) {
    const int i = e % n, j = e / n;
    const int pj = perm[j];
    const double sj = sig[pj];
    ag[i + (int64_t)j * ld] = sj > 0.0 ? su[i + (int64_t)pj * ls] / sj : 0.0;
    vg[i + (int64_t)j * ld] = svv[i + (int64_t)pj * ls];
  }
  for (int j = threadIdx.x; j < n; j += kWG) sv[j] = sig[perm[j]];
}

// n <= 64: one column pair per half-wave (two rows per lane), so the n/2
// disjoint pairs of a round-robin step rotate at the same time; the
// workgroup has 64 * ceil(pairs / 2) threads (up to 1024), so small
// matrices leave room for more workgroups per CU.
// A and V live in LDS (dynamic, 2 n (n+1) doubles, at most 66.5 KB of the
// 160 KB a gfx950 workgroup may use).  Same rotation and stopping rule as
// k_jacobi_svd.
constexpr int kJWG = 1024;
__global__ __launch_bounds__(kJWG) void k_jacobi_svd64(const SmallMat* __restrict__ mats,
                                                       double* __restrict__ A, double* __restrict__ V,
                                                       double* __restrict__ S, int64_t sstride,
                                                       int* __restrict__ sweeps_out, double floor_rel) {
  extern __shared__ double sm[];
  __shared__ double sig[64];
  __shared__ double red[kJWG / 64];
  __shared__ int perm[64];
  __shared__ int rotated;
  const SmallMat m = mats[blockIdx.x];
  const int n = m.n, ld = m.ld;
  double* ag = A + m.off;
  double* vg = V + m.off;
  double* sv = S + (int64_t)blockIdx.x * sstride;
  const int LA = n + 1;
  double* a = sm;
  double* v = sm + n * LA;
  // blockDim.x = 64 * ceil(pairs / 2): one half-wave per column pair
  const int nthr = blockDim.x, nhw = nthr / 32;
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int hl = lane & 31, half = lane >> 5;
  double fro = 0.0;
  for (int e = threadIdx.x; e < n * n; e += nthr) {
    const int i = e % n, j = e / n;
    const double x = ag[i + (int64_t)j * ld];
    fro += x * x;
    a[i + j * LA] = x;
    v[i + j * LA] = (i == j) ? 1.0 : 0.0;
  }
  for (int o = 32; o > 0; o >>= 1) fro += __shfl_xor(fro, o);
  if (lane == 0) red[w] = fro;
  __syncthreads();
  fro = 0.0;
  for (int i = 0; i < nthr / 64; ++i) fro += red[i];
  const double floor2 = floor_rel * fro;
  const int np = n + (n & 1);
  const int circ = np - 1;
  int sweep = 0;
  for (; sweep < 60 && n > 1; ++sweep) {
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int step = 0; step < circ; ++step) {
      for (int pi = w * 2 + half; pi < np / 2; pi += nhw) {
        int p, q;
        if (pi == 0) {
          p = 0;
          q = 1 + step % circ;
        } else {
          p = 1 + (step + pi) % circ;
          q = 1 + (step - pi + circ) % circ;
        }
        if (p >= n || q >= n) continue;
        if (p > q) { const int t = p; p = q; q = t; }
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int i = hl; i < n; i += 32) {
          const double x = a[i + p * LA], y = a[i + q * LA];
          al += x * x;
          be += y * y;
          ga += x * y;
        }
        for (int o = 16; o > 0; o >>= 1) {
          al += __shfl_xor(al, o);
          be += __shfl_xor(be, o);
          ga += __shfl_xor(ga, o);
        }
        if (!(fabs(ga) > 1e-15 * sqrt(al * be)) || !(fabs(ga) > floor2)) continue;
        const double zeta = (be - al) / (2.0 * ga);
        double t;
        if (fabs(zeta) > 1e150)
          t = 0.5 / zeta;
        else
          t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = c * t;
        for (int i = hl; i < n; i += 32) {
          const double x = a[i + p * LA], y = a[i + q * LA];
          a[i + p * LA] = c * x - sn * y;
          a[i + q * LA] = sn * x + c * y;
          const double vx = v[i + p * LA], vy = v[i + q * LA];
          v[i + p * LA] = c * vx - sn * vy;
          v[i + q * LA] = sn * vx + c * vy;
        }
        if (hl == 0) rotated = 1;
      }
      __syncthreads();
    }
    if (!rotated) break;
    __syncthreads();
  }
  if (sweeps_out && threadIdx.x == 0) sweeps_out[blockIdx.x] = sweep;
  for (int j = w * 2 + half; j < n; j += nhw) {
    double s2 = 0.0;
    for (int i = hl; i < n; i += 32) s2 += a[i + j * LA] * a[i + j * LA];
    for (int o = 16; o > 0; o >>= 1) s2 += __shfl_xor(s2, o);
    if (hl == 0) sig[j] = sqrt(s2);
  }
  __syncthreads();
  if ((int)threadIdx.x < n) {  // rank of column j in descending order (ties by index)
    const int j = threadIdx.x;
    int r = 0;
    for (int t = 0; t < n; ++t) r += (sig[t] > sig[j] || (sig[t] == sig[j] && t < j)) ? 1 : 0;
    perm[r] = j;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * n; e += nthr) {
    const int i = e % n, j = e / n;
    const int pj = perm[j];
    const double sj = sig[pj];
    ag[i + (int64_t)j * ld] = sj > 0.0 ? a[i + pj * LA] / sj : 0.0;
    vg[i + (int64_t)j * ld] = v[i + pj * LA];
  }
  for (int j = threadIdx.x; j < n; j += nthr) sv[j] = sig[perm[j]];
}

void jacobi_svd(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, double* A, double* V,
                double* S, int64_t sstride, double* work, hipStream_t s, double floor_rel) {
  if (mats.empty()) return;
  for (auto& mm : mats)
    HMAT_REQUIRE(mm.n <= 512, HMAT_ERR_ARG, "jacobi_svd: n > 512");
  int nmax = 0;
  for (auto& mm : mats) nmax = std::max(nmax, mm.n);
  const size_t lds = (size_t)2 * nmax * (nmax + 1) * sizeof(double);
  const int use_lds = lds <= 128 * 1024 ? 1 : 0;
  dm.upload(mats, s);
  // with profiling on, the sweep counts are recorded as "jacobi_sweeps"
  // (work = total sweeps, launches = calls) and "jacobi_mats"
  DevBuf<int> sw;
  const bool count = prof_enabled();
  if (count) sw.resize(mats.size());
  if (nmax <= 64) {
    // one half-wave per column pair of the round-robin ordering
    const int pairs = (nmax + (nmax & 1)) / 2;
    const unsigned nthr = (unsigned)std::max(64, 64 * ((pairs + 1) / 2));
    k_jacobi_svd64<<<(unsigned)mats.size(), nthr, (size_t)2 * nmax * (nmax + 1) * sizeof(double), s>>>(
        dm.get(), A, V, S, sstride, count ? sw.get() : nullptr, floor_rel);
  } else {
    k_jacobi_svd<<<(unsigned)mats.size(), kWG, use_lds ? lds : 0, s>>>(dm.get(), A, V, S, sstride, work,
                                                                       use_lds, count ? sw.get() : nullptr,
                                                                       floor_rel);
  }
  HIP_CHECK(hipGetLastError());
  if (count) prof_defer_count("jacobi_sweeps", "jacobi_mats", std::move(sw), mats.size(), s);
}

// Gauss-Jordan inversion with partial pivoting, in place.
template <class F>
__global__ __launch_bounds__(kWG) void k_gj_inverse(const SmallMat* __restrict__ mats,
                                                    F* __restrict__ A, int* __restrict__ info) {
  __shared__ F bv[kWG];
  __shared__ int bi[kWG];
  __shared__ int piv_s;
  __shared__ int perm[4096];
  const SmallMat m = mats[blockIdx.x];
  const int n = m.n, ld = m.ld;
  F* a = A + m.off;
  if (threadIdx.x == 0) info[blockIdx.x] = 0;
  for (int k = 0; k < n; ++k) {
    F best = F(-1);
    int bidx = k;
    for (int i = k + threadIdx.x; i < n; i += kWG) {
      const F v = fabs(a[i + (int64_t)k * ld]);
      if (v > best) { best = v; bidx = i; }
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = bidx;
    __syncthreads();
    for (int o = kWG / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        const F v2 = bv[threadIdx.x + o];
        const int i2 = bi[threadIdx.x + o];
        if (v2 > bv[threadIdx.x] || (v2 == bv[threadIdx.x] && i2 < bi[threadIdx.x])) {
          bv[threadIdx.x] = v2;
          bi[threadIdx.x] = i2;
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      piv_s = bi[0];
      perm[k] = bi[0];
      if (!(bv[0] > F(0)) || !isfinite(bv[0])) info[blockIdx.x] = k + 1;
    }
    __syncthreads();
    if (!(bv[0] > F(0)) || !isfinite(bv[0])) return;
    const int p = piv_s;
    if (p != k)
      for (int j = threadIdx.x; j < n; j += kWG) {
        const F t = a[k + (int64_t)j * ld];
        a[k + (int64_t)j * ld] = a[p + (int64_t)j * ld];
        a[p + (int64_t)j * ld] = t;
      }
    __syncthreads();
    const F pv = a[k + (int64_t)k * ld];
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += kWG)
      if (j != k) a[k + (int64_t)j * ld] /= pv;
    __syncthreads();
    for (int e = threadIdx.x; e < n * n; e += kWG) {
      const int i = e % n, j = e / n;
      if (i != k && j != k) a[i + (int64_t)j * ld] -= a[i + (int64_t)k * ld] * a[k + (int64_t)j * ld];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kWG)
      if (i != k) a[i + (int64_t)k * ld] = -a[i + (int64_t)k * ld] / pv;
    if (threadIdx.x == 0) a[k + (int64_t)k * ld] = F(1) / pv;
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const int p = perm[k];
    if (p != k)
      for (int i = threadIdx.x; i < n; i += kWG) {
        const F t = a[i + (int64_t)k * ld];
        a[i + (int64_t)k * ld] = a[i + (int64_t)p * ld];
        a[i + (int64_t)p * ld] = t;
      }
    __syncthreads();
  }
}

// n <= 256: blocked Gauss-Jordan, the same pivoting and breakdown rules as
// k_gj_inverse.  Thread i owns row i.  Per block of NB = 16 columns K:
//   1. the panel A(:, K) is inverted in registers (pivot search by a block
//      argmax, row exchange between the two owning threads, pivot row
//      broadcast through LDS);
//   2. the block's row swaps are applied to the other columns;
//   3. the other columns get the whole block's transformation at once,
//      A(:, c) = [rows not in K] A(:, c) + panel A(K, c), A(K, c) staged in LDS.
// The matrix is read and written once per block instead of once per column.
template <class F, int J>
__device__ __forceinline__ bool gj_panel_col(F (&pan)[16], int kb, int nbk, int n, int lane, int w,
                                             F* bv, int* bi, F* rowk, F* xk, F* xp, int* perm, int* piv_s,
                                             int* bad, int* info) {
  constexpr int NB = 16;
  if (J >= nbk) return true;
  const int i = threadIdx.x;
  const int k = kb + J;
  F v = (i < n && i >= k) ? fabs(pan[J]) : F(-1);
  int idx = i;
  for (int o = 32; o > 0; o >>= 1) {
    const F v2 = __shfl_xor(v, o);
    const int i2 = __shfl_xor(idx, o);
    if (v2 > v || (v2 == v && i2 < idx)) {
      v = v2;
      idx = i2;
    }
  }
  if (lane == 0) {
    bv[w] = v;
    bi[w] = idx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    F best = bv[0];
    int bidx = bi[0];
    for (int q = 1; q < kWG / 64; ++q)
      if (bv[q] > best || (bv[q] == best && bi[q] < bidx)) {
        best = bv[q];
        bidx = bi[q];
      }
    *piv_s = bidx;
    perm[k] = bidx;
    *bad = (!(best > F(0)) || !isfinite(best)) ? 1 : 0;
    if (*bad) info[blockIdx.x] = k + 1;
  }
  __syncthreads();
  if (*bad) return false;
  const int p = *piv_s;
  if (p != k) {  // exchange rows p and k of the panel
    if (i == k)
#pragma unroll
      for (int c = 0; c < NB; ++c) xk[c] = pan[c];
    if (i == p)
#pragma unroll
      for (int c = 0; c < NB; ++c) xp[c] = pan[c];
    __syncthreads();
    if (i == k)
#pragma unroll
      for (int c = 0; c < NB; ++c) pan[c] = xp[c];
    if (i == p)
#pragma unroll
      for (int c = 0; c < NB; ++c) pan[c] = xk[c];
  }
  if (i == k) {
    // one division, then products: 16 serial divisions on this one thread
    // would hold every other thread at the next barrier
    const F rpv = F(1) / pan[J];
    pan[J] = F(1);
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      pan[c] *= rpv;
      rowk[c] = pan[c];
    }
  }
  __syncthreads();
  if (i != k && i < n) {
    const F f = pan[J];
    pan[J] = F(0);
#pragma unroll
    for (int c = 0; c < NB; ++c) pan[c] -= f * rowk[c];
  }
  return true;
}

template <class F, int J>
__device__ __forceinline__ bool gj_panel(F (&pan)[16], int kb, int nbk, int n, int lane, int w, F* bv,
                                         int* bi, F* rowk, F* xk, F* xp, int* perm, int* piv_s, int* bad,
                                         int* info) {
  if constexpr (J < 16) {
    if (!gj_panel_col<F, J>(pan, kb, nbk, n, lane, w, bv, bi, rowk, xk, xp, perm, piv_s, bad, info))
      return false;
    return gj_panel<F, J + 1>(pan, kb, nbk, n, lane, w, bv, bi, rowk, xk, xp, perm, piv_s, bad, info);
  }
  return true;
}

template <class F>
__global__ __launch_bounds__(kWG) void k_gj_blocked(const SmallMat* __restrict__ mats, F* __restrict__ A,
                                                    int* __restrict__ info) {
  constexpr int NB = 16;
  __shared__ F R[NB * 256];  // R[j * 256 + c] = A(kb + j, c) after the block's swaps
  __shared__ F rowk[NB], xk[NB], xp[NB];
  __shared__ F bv[kWG / 64];
  __shared__ int bi[kWG / 64];
  __shared__ int perm[256];
  __shared__ int piv_s, bad;
  const SmallMat m = mats[blockIdx.x];
  const int n = m.n, ld = m.ld;
  F* a = A + m.off;
  const int i = threadIdx.x, lane = threadIdx.x % 64, w = threadIdx.x / 64;
  if (i == 0) info[blockIdx.x] = 0;
  for (int kb = 0; kb < n; kb += NB) {
    const int nbk = min(NB, n - kb);
    F pan[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) pan[j] = (i < n && j < nbk) ? a[i + (int64_t)(kb + j) * ld] : F(0);
    if (!gj_panel<F, 0>(pan, kb, nbk, n, lane, w, bv, bi, rowk, xk, xp, perm, &piv_s, &bad, info))
      return;  // info holds the failing column + 1
    // 2. this block's row swaps on the other columns (one thread per column)
    if (i < n && (i < kb || i >= kb + nbk))
      for (int j = 0; j < nbk; ++j) {
        const int k = kb + j, p = perm[k];
        if (p != k) {
          const F t = a[k + (int64_t)i * ld];
          a[k + (int64_t)i * ld] = a[p + (int64_t)i * ld];
          a[p + (int64_t)i * ld] = t;
        }
      }
    __syncthreads();
    // rows of R beyond nbk are zero (the update runs over all NB rows)
    if (i < n && (i < kb || i >= kb + nbk))
      for (int j = 0; j < NB; ++j) R[j * 256 + i] = j < nbk ? a[kb + j + (int64_t)i * ld] : F(0);
    __syncthreads();
    // 3. other columns c: A(i, c) = [i not in K] A(i, c) + sum_j M(i, j) R(j, c)
    //    (the pivot rows are replaced, not updated: forming M - I there would
    //    cancel badly when the pivots are large)
    if (i < n) {
      const bool in_k = i >= kb && i < kb + nbk;
      F wv[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) wv[j] = (j < nbk) ? pan[j] : F(0);
      for (int c = 0; c < n; ++c) {
        if (c >= kb && c < kb + nbk) continue;
        F acc = in_k ? F(0) : a[i + (int64_t)c * ld];
#pragma unroll
        for (int j = 0; j < NB; ++j) acc = fma(wv[j], R[j * 256 + c], acc);
        a[i + (int64_t)c * ld] = acc;
      }
#pragma unroll
      for (int j = 0; j < NB; ++j)
        if (j < nbk) a[i + (int64_t)(kb + j) * ld] = pan[j];
    }
    __syncthreads();
  }
  // column swaps in reverse order (each thread its own row)
  if (i < n)
    for (int k = n - 1; k >= 0; --k) {
      const int p = perm[k];
      if (p != k) {
        const F t = a[i + (int64_t)k * ld];
        a[i + (int64_t)k * ld] = a[i + (int64_t)p * ld];
        a[i + (int64_t)p * ld] = t;
      }
    }
}

template <class F>
void gj_inverse(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, F* A, int* info,
                hipStream_t s) {
  if (mats.empty()) return;
  for (auto& mm : mats) HMAT_REQUIRE(mm.n <= 4096, HMAT_ERR_ARG, "gj_inverse: n > 4096");
  int nmax = 0;
  for (auto& mm : mats) nmax = std::max(nmax, mm.n);
  dm.upload(mats, s);
  if (nmax <= 256)
    k_gj_blocked<F><<<(unsigned)mats.size(), kWG, 0, s>>>(dm.get(), A, info);
  else
    k_gj_inverse<F><<<(unsigned)mats.size(), kWG, 0, s>>>(dm.get(), A, info);
  HIP_CHECK(hipGetLastError());
}
template void gj_inverse<double>(const std::vector<SmallMat>&, DevBuf<SmallMat>&, double*, int*, hipStream_t);
template void gj_inverse<float>(const std::vector<SmallMat>&, DevBuf<SmallMat>&, float*, int*, hipStream_t);

// --------------------------------------------------------------------------
// Fill helpers
// --------------------------------------------------------------------------
template <class F>
__global__ void k_fill_identity(const SegNN* __restrict__ segs, F* __restrict__ Y, int64_t ldy) {
  const SegNN sg = segs[blockIdx.y];
  const int64_t total = (int64_t)sg.len * sg.q;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % sg.len), c = (int)(e / sg.len);
    Y[sg.yoff + r + (int64_t)c * ldy] = (r == c) ? F(1) : F(0);
  }
}

template <class F>
void fill_identity_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, F* Y, int64_t ldy,
                        hipStream_t s) {
  if (segs.empty()) return;
  dsegs.upload(segs, s);
  k_fill_identity<F><<<dim3(64, (unsigned)segs.size()), kWG, 0, s>>>(dsegs.get(), Y, ldy);
  HIP_CHECK(hipGetLastError());
}
template void fill_identity_segs<double>(const std::vector<SegNN>&, DevBuf<SegNN>&, double*, int64_t, hipStream_t);
template void fill_identity_segs<float>(const std::vector<SegNN>&, DevBuf<SegNN>&, float*, int64_t, hipStream_t);

template <class To, class Ti>
__global__ void k_convert(const Ti* __restrict__ X, To* __restrict__ Y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    Y[i] = (To)X[i];
}
template <class To, class Ti>
void convert(const Ti* X, To* Y, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>(ceil_div(n, kWG), 4096);
  k_convert<To, Ti><<<g, kWG, 0, s>>>(X, Y, n);
  HIP_CHECK(hipGetLastError());
}
template void convert<float, double>(const double*, float*, int64_t, hipStream_t);
template void convert<double, float>(const float*, double*, int64_t, hipStream_t);

template <class To, class Ti>
__global__ void k_convert_cols(int64_t n, const Ti* __restrict__ X, int64_t ldx, To* __restrict__ Y,
                               int64_t ldy) {
  const int c = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    Y[i + c * ldy] = (To)X[i + c * ldx];
}
template <class To, class Ti>
void convert_cols(int64_t n, int ncols, const Ti* X, int64_t ldx, To* Y, int64_t ldy, hipStream_t s) {
  if (n <= 0 || ncols <= 0) return;
  const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(n, kWG), 1024);
  k_convert_cols<To, Ti><<<dim3(gx, (unsigned)ncols), kWG, 0, s>>>(n, X, ldx, Y, ldy);
  HIP_CHECK(hipGetLastError());
}
template void convert_cols<float, double>(int64_t, int, const double*, int64_t, float*, int64_t, hipStream_t);
template void convert_cols<double, float>(int64_t, int, const float*, int64_t, double*, int64_t, hipStream_t);
template void convert_cols<double, double>(int64_t, int, const double*, int64_t, double*, int64_t, hipStream_t);

__global__ void k_add_f32_cols(int64_t n, const float* __restrict__ D, int64_t ldd,
                               double* __restrict__ X, int64_t ldx) {
  const int c = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    X[i + c * ldx] += (double)D[i + c * ldd];
}
void add_f32_cols(int64_t n, int ncols, const float* D, int64_t ldd, double* X, int64_t ldx,
                  hipStream_t s) {
  if (n <= 0 || ncols <= 0) return;
  const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(n, kWG), 1024);
  k_add_f32_cols<<<dim3(gx, (unsigned)ncols), kWG, 0, s>>>(n, D, ldd, X, ldx);
  HIP_CHECK(hipGetLastError());
}

__global__ void k_copy_segs(const SegNN* __restrict__ segs, const double* __restrict__ X, int64_t ldx,
                            double* __restrict__ Y, int64_t ldy) {
  const SegNN sg = segs[blockIdx.y];
  const int64_t total = (int64_t)sg.len * sg.q;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % sg.len), c = (int)(e / sg.len);
    Y[sg.yoff + r + (int64_t)c * ldy] = X[sg.xoff + r + (int64_t)c * ldx];
  }
}

__global__ void k_transpose_blocks(double* __restrict__ base, int64_t ld, const int32_t* __restrict__ lens) {
  const int n = lens[blockIdx.x];
  double* a = base + (int64_t)blockIdx.x * ld * ld;
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e % n, j = e / n;
    if (i < j) {  // each pair once
      const double t = a[i + (int64_t)j * ld];
      a[i + (int64_t)j * ld] = a[j + (int64_t)i * ld];
      a[j + (int64_t)i * ld] = t;
    }
  }
}

void transpose_blocks(double* base, int64_t ld, const std::vector<int32_t>& lens, hipStream_t s) {
  if (lens.empty()) return;
  k_transpose_blocks<<<(unsigned)lens.size(), kWG, 0, s>>>(base, ld, stage(lens, s));
  HIP_CHECK(hipGetLastError());
}

void copy_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, const double* X, int64_t ldx,
               double* Y, int64_t ldy, hipStream_t s) {
  if (segs.empty()) return;
  dsegs.upload(segs, s);
  k_copy_segs<<<dim3(64, (unsigned)segs.size()), kWG, 0, s>>>(dsegs.get(), X, ldx, Y, ldy);
  HIP_CHECK(hipGetLastError());
}

__global__ void k_trunc_prep(const SmallMat* __restrict__ mats, const double* __restrict__ Ur,
                             const double* __restrict__ Vr, const double* __restrict__ S,
                             int64_t sstride, const int* __restrict__ rank, int rmax,
                             double* __restrict__ TU, double* __restrict__ TV, int s_on_ur) {
  const SmallMat m = mats[blockIdx.x];
  const int r = rank[blockIdx.x];
  const double* sv = S + (int64_t)blockIdx.x * sstride;
  for (int e = threadIdx.x; e < m.n * rmax; e += blockDim.x) {
    const int i = e % m.n, j = e / m.n;
    const int64_t o = m.off + i + (int64_t)j * m.ld;
    const double sj = j < r ? sv[j] : 0.0;
    TU[o] = j < r ? (s_on_ur ? Vr[o] : Vr[o] * sj) : 0.0;
    TV[o] = j < r ? (s_on_ur ? Ur[o] * sj : Ur[o]) : 0.0;
  }
}

void trunc_prep(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, const double* Ur,
                const double* Vr, const double* S, int64_t sstride, const int* rank, int rmax,
                double* TU, double* TV, int s_on_ur, hipStream_t s) {
  if (mats.empty()) return;
  dm.upload(mats, s);
  k_trunc_prep<<<(unsigned)mats.size(), kWG, 0, s>>>(dm.get(), Ur, Vr, S, sstride, rank, rmax, TU, TV,
                                                     s_on_ur);
  HIP_CHECK(hipGetLastError());
}

// Batched small GEMMs, one workgroup per item: C = alpha op(A) op(B) +
// beta C in output tiles of 64 x 32, K in chunks of 32.  Both operands are
// loaded with consecutive lanes on consecutive addresses (either transpose)
// and stored k-major in LDS; each thread then accumulates a 4 x 2 block of
// the tile from LDS.  A C that aliases A or B is not supported.
constexpr int kBgTM = 64, kBgTN = 32, kBgKC = 32;
__global__ __launch_bounds__(kWG) void k_bgemm(const BGemm* __restrict__ items) {
  __shared__ double As[kBgKC][kBgTM + 1];
  __shared__ double Bs[kBgKC][kBgTN + 1];
  const BGemm g = items[blockIdx.x];
  const int tid = threadIdx.x;
  const int ti = tid % 16, tj = tid / 16;  // rows 4 ti .. 4 ti + 3, columns 2 tj, 2 tj + 1
  for (int i0 = 0; i0 < g.m; i0 += kBgTM)
    for (int j0 = 0; j0 < g.n; j0 += kBgTN) {
      double acc[4][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
      for (int l0 = 0; l0 < g.k; l0 += kBgKC) {
        // op(A)(i0 + i, l0 + l) -> As[l][i]
#pragma unroll
        for (int q = 0; q < kBgTM * kBgKC / kWG; ++q) {
          const int e = tid + q * kWG;
          int i, l;
          if (g.ta) {
            l = e % kBgKC;
            i = e / kBgKC;
          } else {
            i = e % kBgTM;
            l = e / kBgTM;
          }
          const int gi = i0 + i, gl = l0 + l;
          double v = 0.0;
          if (gi < g.m && gl < g.k) v = g.ta ? g.A[gl + (int64_t)gi * g.lda] : g.A[gi + (int64_t)gl * g.lda];
          As[l][i] = v;
        }
        // op(B)(l0 + l, j0 + j) -> Bs[l][j]
#pragma unroll
        for (int q = 0; q < kBgTN * kBgKC / kWG; ++q) {
          const int e = tid + q * kWG;
          int l, j;
          if (g.tb) {
            j = e % kBgTN;
            l = e / kBgTN;
          } else {
            l = e % kBgKC;
            j = e / kBgKC;
          }
          const int gl = l0 + l, gj = j0 + j;
          double v = 0.0;
          if (gl < g.k && gj < g.n) v = g.tb ? g.B[gj + (int64_t)gl * g.ldb] : g.B[gl + (int64_t)gj * g.ldb];
          Bs[l][j] = v;
        }
        __syncthreads();
        const int lmax = min(kBgKC, g.k - l0);
        for (int l = 0; l < lmax; ++l) {
          double a[4], b[2];
#pragma unroll
          for (int r = 0; r < 4; ++r) a[r] = As[l][4 * ti + r];
          b[0] = Bs[l][2 * tj];
          b[1] = Bs[l][2 * tj + 1];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            acc[r][0] = fma(a[r], b[0], acc[r][0]);
            acc[r][1] = fma(a[r], b[1], acc[r][1]);
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int gi = i0 + 4 * ti + r, gj = j0 + 2 * tj + c;
          if (gi < g.m && gj < g.n) {
            double* cp = g.C + gi + (int64_t)gj * g.ldc;
            *cp = g.alpha * acc[r][c] + (g.beta != 0.0 ? g.beta * *cp : 0.0);
          }
        }
    }
}

void bgemm(const std::vector<BGemm>& items, DevBuf<BGemm>& d, hipStream_t s) {
  if (items.empty()) return;
  d.upload(items, s);
  k_bgemm<<<(unsigned)items.size(), kWG, 0, s>>>(d.get());
  HIP_CHECK(hipGetLastError());
}

// Node tasks: one workgroup runs the ops of one task in order (per-node
// work of the HSS sweeps, where each op reads what the previous one wrote,
// so an op's cost is its latency).  An op is a small GEMM (m, k up to a few
// hundred, n = right-hand sides): op(B) for up to kNtK reduction indices and
// eight columns is staged in LDS (one load per thread, one barrier); each
// thread owns one output row, loads its op(A) entries 16 at a time ahead of
// the FMAs (coalesced down the rows for ta = 0; the small generator
// matrices sit in L2; the first batch is issued before op(B) is staged, so
// the two loads overlap), and the reduction is split over 256 / R thread
// groups (R = 64, 128 or 256 rows per pass) and summed through LDS in group
// order.
constexpr int kNtCols = 8, kNtK = 512, kNtA = 16;
__global__ __launch_bounds__(kWG) void k_node_tasks(const NodeTask* __restrict__ tasks,
                                                    const BGemm* __restrict__ ops) {
  __shared__ double red[kWG * kNtCols];
  __shared__ double bs[kNtK * kNtCols];  // op(B)(l, j0 + jj) at bs[l * kNtCols + jj]
  const NodeTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x;
  for (int o = 0; o < t.nops; ++o) {
    const BGemm g = ops[t.op0 + o];
    if (g.k == 0) {
      // copy: C = alpha A + beta C (A m x n, leading dimension lda)
      for (int e = tid; e < g.m * g.n; e += kWG) {
        const int i = e % g.m, j = e / g.m;
        double* c = g.C + i + (int64_t)j * g.ldc;
        const double v = g.alpha * g.A[i + (int64_t)j * g.lda];
        *c = g.beta != 0.0 ? v + g.beta * *c : v;
      }
      __syncthreads();  // the next op may read this op's C
      continue;
    }
    const int R = g.m <= 64 ? 64 : (g.m <= 128 ? 128 : 256);
    const int ns = kWG / R;
    const int r = tid % R, sp = tid / R;
    const int64_t as = g.ta ? 1 : g.lda;  // step of l in op(A)(i, l)
    const int64_t bl = g.tb ? g.ldb : 1;  // step of l in op(B)(l, j)
    const int64_t bj = g.tb ? 1 : g.ldb;  // step of j
    for (int j0 = 0; j0 < g.n; j0 += kNtCols) {
      const int nj = min(kNtCols, g.n - j0);
      for (int i0 = 0; i0 < g.m; i0 += R) {
        const int i = i0 + r;
        const double* ap = g.A + (g.ta ? (int64_t)i * g.lda : i);
        double acc[kNtCols];
#pragma unroll
        for (int jj = 0; jj < kNtCols; ++jj) acc[jj] = 0.0;
        for (int k0 = 0; k0 < g.k; k0 += kNtK) {
          const int kb = min(kNtK, g.k - k0);
          const int kc = (kb + ns - 1) / ns;
          const int l0 = min(kb, sp * kc), l1 = min(kb, l0 + kc);
          const bool row = i < g.m;
          double a[kNtA];
#pragma unroll
          for (int q = 0; q < kNtA; ++q) a[q] = (row && l0 + q < l1) ? ap[(int64_t)(k0 + l0 + q) * as] : 0.0;
          for (int e = tid; e < kb * kNtCols; e += kWG) {
            const int l = e / kNtCols, jj = e % kNtCols;
            bs[e] = jj < nj ? g.B[(int64_t)(k0 + l) * bl + (int64_t)(j0 + jj) * bj] : 0.0;
          }
          __syncthreads();
          if (row)
            for (int lb = l0; lb < l1; lb += kNtA) {
              if (lb > l0) {
#pragma unroll
                for (int q = 0; q < kNtA; ++q) a[q] = lb + q < l1 ? ap[(int64_t)(k0 + lb + q) * as] : 0.0;
              }
#pragma unroll
              for (int q = 0; q < kNtA; ++q) {
                if (lb + q >= l1) break;
                const double* b = bs + (lb + q) * kNtCols;
#pragma unroll
                for (int jj = 0; jj < kNtCols; ++jj) acc[jj] = fma(a[q], b[jj], acc[jj]);
              }
            }
          __syncthreads();  // bs is rewritten next
        }
        if (ns > 1) {
#pragma unroll
          for (int jj = 0; jj < kNtCols; ++jj) red[(sp * R + r) * kNtCols + jj] = acc[jj];
          __syncthreads();
          if (sp == 0)
            for (int q = 1; q < ns; ++q)
#pragma unroll
              for (int jj = 0; jj < kNtCols; ++jj) acc[jj] += red[(q * R + r) * kNtCols + jj];
          __syncthreads();
        }
        if (sp == 0 && i < g.m)
          for (int jj = 0; jj < nj; ++jj) {
            double* c = g.C + i + (int64_t)(j0 + jj) * g.ldc;
            const double v = g.alpha * acc[jj];
            *c = g.beta != 0.0 ? v + g.beta * *c : v;
          }
      }
    }
    __syncthreads();  // the next op may read this op's C
  }
}

void node_tasks(const NodeTask* dtasks, int ntasks, const BGemm* dops, hipStream_t s) {
  if (ntasks <= 0) return;
  k_node_tasks<<<(unsigned)ntasks, kWG, 0, s>>>(dtasks, dops);
  HIP_CHECK(hipGetLastError());
}

__global__ void k_rect_axpby(const Rect* __restrict__ items) {
  const Rect r = items[blockIdx.y];
  const int64_t total = (int64_t)r.m * r.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % r.m), j = (int)(e / r.m);
    double* d = r.dst + i + (int64_t)j * r.ldd;
    const double sv = r.src ? r.src[i + (int64_t)j * r.lds] : 0.0;
    *d = r.alpha * sv + (r.beta != 0.0 ? r.beta * *d : 0.0);
  }
}
__global__ void k_rect_add_identity(const Rect* __restrict__ items) {
  const Rect r = items[blockIdx.y];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < min(r.m, r.n); i += gridDim.x * blockDim.x)
    r.dst[i + (int64_t)i * r.ldd] += r.alpha;
}
void rect_axpby(const std::vector<Rect>& items, DevBuf<Rect>& d, hipStream_t s) {
  if (items.empty()) return;
  d.upload(items, s);
  k_rect_axpby<<<dim3(16, (unsigned)items.size()), kWG, 0, s>>>(d.get());
  HIP_CHECK(hipGetLastError());
}
void rect_add_identity(const std::vector<Rect>& items, DevBuf<Rect>& d, hipStream_t s) {
  if (items.empty()) return;
  d.upload(items, s);
  k_rect_add_identity<<<dim3(4, (unsigned)items.size()), kWG, 0, s>>>(d.get());
  HIP_CHECK(hipGetLastError());
}
// dst(j, i) = alpha src(i, j): reads run down the source columns (small
// blocks; the strided writes stay within a few KB)
__global__ void k_rect_transpose(const Rect* __restrict__ items) {
  const Rect r = items[blockIdx.y];
  const int64_t total = (int64_t)r.m * r.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % r.m), j = (int)(e / r.m);
    double* d = r.dst + j + (int64_t)i * r.ldd;
    const double sv = r.src[i + (int64_t)j * r.lds];
    *d = r.alpha * sv + (r.beta != 0.0 ? r.beta * *d : 0.0);
  }
}
void rect_transpose(const std::vector<Rect>& items, DevBuf<Rect>& d, hipStream_t s) {
  if (items.empty()) return;
  d.upload(items, s);
  k_rect_transpose<<<dim3(16, (unsigned)items.size()), kWG, 0, s>>>(d.get());
  HIP_CHECK(hipGetLastError());
}

__global__ void k_zero_rects(const ZeroRect* __restrict__ items) {
  const ZeroRect z = items[blockIdx.y];
  const int64_t total = (int64_t)z.len * z.ncols;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
    z.p[(z.r0 + e % z.len) + (z.c0 + e / z.len) * z.ld] = 0.0;
}

// Byte fills and copies in the widest unit the addresses and size allow.
template <class U>
__global__ void k_fill_units(U* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = U{};
}
template <class U>
__global__ void k_copy_units(U* __restrict__ d, const U* __restrict__ src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = src[i];
}
static unsigned unit_grid(int64_t n) { return (unsigned)std::min<int64_t>(std::max<int64_t>((n + kWG - 1) / kWG, 1), 4096); }

}  // namespace kern

void zero_bytes_async(void* p, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  const uintptr_t a = reinterpret_cast<uintptr_t>(p) | bytes;
  if ((a & 15) == 0) {
    const int64_t n = (int64_t)(bytes / 16);
    kern::k_fill_units<uint4><<<kern::unit_grid(n), kern::kWG, 0, s>>>(static_cast<uint4*>(p), n);
  } else if ((a & 3) == 0) {
    const int64_t n = (int64_t)(bytes / 4);
    kern::k_fill_units<uint32_t><<<kern::unit_grid(n), kern::kWG, 0, s>>>(static_cast<uint32_t*>(p), n);
  } else {
    const int64_t n = (int64_t)bytes;
    kern::k_fill_units<uint8_t><<<kern::unit_grid(n), kern::kWG, 0, s>>>(static_cast<uint8_t*>(p), n);
  }
  HIP_CHECK(hipGetLastError());
}

void copy_bytes_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  const uintptr_t a = reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | bytes;
  if ((a & 15) == 0) {
    const int64_t n = (int64_t)(bytes / 16);
    kern::k_copy_units<uint4><<<kern::unit_grid(n), kern::kWG, 0, s>>>(static_cast<uint4*>(dst),
                                                                        static_cast<const uint4*>(src), n);
  } else if ((a & 3) == 0) {
    const int64_t n = (int64_t)(bytes / 4);
    kern::k_copy_units<uint32_t><<<kern::unit_grid(n), kern::kWG, 0, s>>>(
        static_cast<uint32_t*>(dst), static_cast<const uint32_t*>(src), n);
  } else {
    const int64_t n = (int64_t)bytes;
    kern::k_copy_units<uint8_t><<<kern::unit_grid(n), kern::kWG, 0, s>>>(static_cast<uint8_t*>(dst),
                                                                          static_cast<const uint8_t*>(src), n);
  }
  HIP_CHECK(hipGetLastError());
}

namespace kern {

// D_i(j, j) += *shift for j < len_i over every leaf block (lmax x lmax each).
__global__ void k_add_diag(double* __restrict__ D, int64_t lm, const int* __restrict__ lens, int nleaf,
                           const double* __restrict__ shift) {
  const double v = *shift;
  const int64_t total = (int64_t)nleaf * lm;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / lm), j = (int)(e % lm);
    if (j < lens[i]) D[(int64_t)i * lm * lm + (int64_t)j * lm + j] += v;
  }
}

__global__ void k_set_scalar(double* p, double v) {
  if (threadIdx.x == 0) p[threadIdx.x] = v;
}

void set_scalar(double* p, double v, hipStream_t s) {
  k_set_scalar<<<1, 64, 0, s>>>(p, v);
  HIP_CHECK(hipGetLastError());
}

void add_diag(double* D, int64_t lm, const int* dlens, int nleaf, const double* dshift, hipStream_t s) {
  const int64_t n = (int64_t)nleaf * lm;
  if (n == 0) return;
  k_add_diag<<<unit_grid(n), kWG, 0, s>>>(D, lm, dlens, nleaf, dshift);
  HIP_CHECK(hipGetLastError());
}

void zero_rects(const std::vector<ZeroRect>& items, DevBuf<ZeroRect>& d, hipStream_t s) {
  if (items.empty()) return;
  d.upload(items, s);
  k_zero_rects<<<dim3(32, (unsigned)items.size()), kWG, 0, s>>>(d.get());
  HIP_CHECK(hipGetLastError());
}

__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double u01(uint64_t z) { return (double)(z >> 11) * (1.0 / 9007199254740992.0); }

// Omega entry (r, c) of stream `id` (the segment's toff field).
__global__ void k_fill_random(const SegNN* __restrict__ segs, double* __restrict__ Y, int64_t ldy,
                              uint64_t seed, int dist) {
  const SegNN sg = segs[blockIdx.y];
  const uint64_t key = splitmix64(seed ^ splitmix64((uint64_t)sg.toff + 0x51ED2701ull));
  const int64_t total = (int64_t)sg.len * sg.q;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e % sg.len), c = (int)(e / sg.len);
    const uint64_t ctr = ((uint64_t)c << 40) ^ (uint64_t)r;
    double val;
    if (dist == HMAT_OMEGA_GAUSSIAN) {
      const double u1 = u01(splitmix64(key ^ (ctr * 2 + 0x1234567ull)));
      const double u2 = u01(splitmix64(key ^ (ctr * 2 + 1 + 0x7654321ull)));
      val = sqrt(-2.0 * log(1.0 - u1)) * cospi(2.0 * u2);
    } else {
      val = u01(splitmix64(key ^ ctr));
    }
    Y[sg.yoff + r + (int64_t)c * ldy] = val;
  }
}

void fill_random_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, double* Y,
                      int64_t ldy, uint64_t seed, int dist, hipStream_t s) {
  if (segs.empty()) return;
  dsegs.upload(segs, s);
  k_fill_random<<<dim3(64, (unsigned)segs.size()), kWG, 0, s>>>(dsegs.get(), Y, ldy, seed, dist);
  HIP_CHECK(hipGetLastError());
}

template <class Gen>
__global__ void k_gen_blocks(Gen gen, const BlockDesc* __restrict__ blocks, double* __restrict__ out) {
  const BlockDesc b = blocks[blockIdx.y];
  const int64_t total = (int64_t)b.m * b.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % b.m), j = (int)(e / b.m);
    const int64_t gi = b.r0 + i, gj = b.c0 + j;
    out[b.off + i + (int64_t)j * b.ld] = gen.eval(gen.row_pt(gi), gen.col_pt(gj), gi, gj);
  }
}

void gen_blocks(const GenDesc& g, const std::vector<BlockDesc>& blocks, DevBuf<BlockDesc>& db,
                double* out, hipStream_t s) {
  if (blocks.empty()) return;
  db.upload(blocks, s);
  with_gen(g, [&](auto gen) {
    k_gen_blocks<<<dim3(64, (unsigned)blocks.size()), kWG, 0, s>>>(gen, db.get(), out);
  });
  HIP_CHECK(hipGetLastError());
}

__global__ void k_axpby_cols(int64_t n, double alpha, const double* __restrict__ X, int64_t ldx,
                             double beta, const double* __restrict__ Y, int64_t ldy,
                             double* __restrict__ Z, int64_t ldz) {
  const int c = blockIdx.y;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    Z[i + c * ldz] = alpha * X[i + c * ldx] + beta * Y[i + c * ldy];
}

void axpby_cols(int64_t n, int ncols, double alpha, const double* X, int64_t ldx, double beta,
                const double* Y, int64_t ldy, double* Z, int64_t ldz, hipStream_t s) {
  if (n <= 0 || ncols <= 0) return;
  const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(n, kWG), 1024);
  k_axpby_cols<<<dim3(gx, (unsigned)ncols), kWG, 0, s>>>(n, alpha, X, ldx, beta, Y, ldy, Z, ldz);
  HIP_CHECK(hipGetLastError());
}

// Squared column norms in two deterministic stages: kNormParts workgroups
// per column sum contiguous chunks, then one workgroup per column adds the
// partial sums in order.
constexpr int kNormParts = 64;
__global__ void k_col_norms2_part(const double* __restrict__ X, int64_t ldx, int64_t n,
                                  double* __restrict__ part) {
  __shared__ double red[kWG / 64];
  const double* x = X + (int64_t)blockIdx.y * ldx;
  const int64_t chunk = (n + kNormParts - 1) / kNormParts;
  const int64_t b = blockIdx.x * chunk, e = min(n, b + chunk);
  double s = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += kWG) s += x[i] * x[i];
  s = block_sum(s, red);
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * kNormParts + blockIdx.x] = s;
}
__global__ void k_col_norms2_sum(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[kWG / 64];
  double s = threadIdx.x < kNormParts ? part[(int64_t)blockIdx.x * kNormParts + threadIdx.x] : 0.0;
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

void col_norms2(const double* X, int64_t ldx, int64_t n, int ncols, double* out, double* part, hipStream_t s) {
  if (ncols <= 0) return;
  k_col_norms2_part<<<dim3(kNormParts, (unsigned)ncols), kWG, 0, s>>>(X, ldx, n, part);
  HIP_CHECK(hipGetLastError());
  k_col_norms2_sum<<<(unsigned)ncols, kWG, 0, s>>>(part, out);
  HIP_CHECK(hipGetLastError());
}

}  // namespace kern
}  // namespace hmat
