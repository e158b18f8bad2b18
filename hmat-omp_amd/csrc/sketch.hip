// Block sketch GEMMs of the build: Y = A(rows, red) X or A(red, rows)^T X
// for every off-diagonal block of a tree depth in one launch, with A's
// entries generated in registers (or read from a dense input).
//
//   k_sketch     FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 entries.
//   k_sketch_lp  BF16 MFMA (v_mfma_f32_16x16x32_bf16): FP32 entries split
//                into SPLIT BF16 terms, X pre-split into BF16 planes,
//                products whose term indices sum to < SPLIT, FP32
//                accumulation flushed to FP64 every 512 reduction indices.
//
// Both stage a K-tile of 64 reduction rows of X (and their points) in LDS
// and zero the X rows of reduction indices past the end or inside the
// excluded interval as they stage them, so their inner loops carry no
// masks.  The next tile is loaded into registers during the current one's
// products; the multi-term k_sketch_lp writes it to a second LDS buffer, so
// a tile costs it one barrier.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "generators.hpp"
#include "kernels.hpp"
#include "mfma.hpp"

namespace hmat {
namespace kern {

constexpr int kWGs = 256;
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

struct SkTile {
  int32_t seg, row0, rb, re;
  int32_t split, pad_;
};

// rows [kb, kb+64) of the reduction range are all in range and outside the
// excluded interval
__device__ inline bool ktile_full(const SegSk& sg, int kb, int re) {
  if (kb + 64 > re) return false;
  if (sg.excllen <= 0) return true;
  const int64_t lo = sg.red0 + kb, hi = lo + 64;
  return hi <= sg.excl0 || lo >= sg.excl0 + sg.excllen;
}

template <class Gen, bool TRANS, int NT, int RF>
__global__ __launch_bounds__(kWGs) void k_sketch(Gen gen, const SegSk* __restrict__ segs,
                                                 const SkTile* __restrict__ tiles,
                                                 const double* __restrict__ X, int64_t ldx,
                                                 double* __restrict__ Y, int64_t ldy,
                                                 int64_t split_stride) {
  constexpr int KT = 64, NC = NT * 16;
  __shared__ double xs[KT][NC + 1];
  __shared__ double rp[KT];
  const SkTile tl = tiles[blockIdx.x];
  const SegSk sg = segs[tl.seg];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int cbase = blockIdx.y * NC;
  const int kcols = min(NC, sg.k - cbase);
  if (kcols <= 0) return;  // a column group past this segment's width: nothing to store
  mfma::v4d acc[RF][NT];
#pragma unroll
  for (int f = 0; f < RF; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[f][t] = mfma::v4d{0.0, 0.0, 0.0, 0.0};
  // Rows past the block's end take the last row's point (finite entries)
  // and are never stored.  Reduction indices past the end or inside the
  // excluded interval take the last index's point and meet zero rows of X.
  int64_t iglob[RF];
  double opt[RF];
#pragma unroll
  for (int f = 0; f < RF; ++f) {
    const int rloc = tl.row0 + w * 16 * RF + f * 16 + (lane % 16);
    iglob[f] = sg.out0 + min(rloc, max(sg.outlen - 1, 0));
    opt[f] = TRANS ? gen.col_pt(iglob[f]) : gen.row_pt(iglob[f]);
  }
  const int kq = lane / 16;
  // Software pipeline: the next K-tile of X and its points are loaded into
  // registers while the current tile is computed from LDS.
  constexpr int NPRE = KT * NC / kWGs;
  static_assert(KT * NC % kWGs == 0, "prefetch layout");
  double pre[NPRE];
  double ppt = 0.0;
  auto issue = [&](int kb) {
#pragma unroll
    for (int i = 0; i < NPRE; ++i) {
      const int e = threadIdx.x + i * kWGs;
      const int kk = e % KT, cc = e / KT, j = kb + kk;
      const bool excl = (uint64_t)(sg.red0 + j - sg.excl0) < (uint64_t)sg.excllen;
      // every load in bounds (clamped index), masked by a select: no branches
      const double v = X[sg.xoff + min(j, tl.re - 1) + (int64_t)(cbase + min(cc, kcols - 1)) * ldx];
      pre[i] = (j < tl.re && cc < kcols && !excl) ? v : 0.0;
    }
    if (threadIdx.x < KT) {
      const int j = kb + threadIdx.x;
      ppt = TRANS ? gen.row_pt(sg.red0 + min(j, tl.re - 1)) : gen.col_pt(sg.red0 + min(j, tl.re - 1));
    }
  };
  auto commit = [&]() {
#pragma unroll
    for (int i = 0; i < NPRE; ++i) {
      const int e = threadIdx.x + i * kWGs;
      xs[e % KT][e / KT] = pre[i];
    }
    if (threadIdx.x < KT) rp[threadIdx.x] = ppt;
  };
  // Dense input reads A from memory in the inner loop already; there the
  // next tile is fetched after the compute, keeping its registers free.
  constexpr bool PF = !std::is_same<Gen, GDense>::value;
  issue(tl.rb);
  commit();
  __syncthreads();
  for (int kb = tl.rb; kb < tl.re; kb += KT) {
    const bool has_next = kb + KT < tl.re;
    if (PF && has_next) issue(kb + KT);
#pragma unroll 4
    for (int ks = 0; ks < KT / 4; ++ks) {
      const int kk = ks * 4 + kq;
      const int jl = min(kb + kk, tl.re - 1);
      const int64_t jg = sg.red0 + jl;
      const double rpt = rp[kk];
      double a[RF];
#pragma unroll
      for (int f = 0; f < RF; ++f)
        a[f] = TRANS ? gen.eval(rpt, opt[f], jg, iglob[f]) : gen.eval(opt[f], rpt, iglob[f], jg);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const double b = xs[kk][t * 16 + (lane % 16)];
#pragma unroll
        for (int f = 0; f < RF; ++f) acc[f][t] = mfma::f64_16x16x4(a[f], b, acc[f][t]);
      }
    }
    __syncthreads();
    if (has_next) {
      if (!PF) issue(kb + KT);
      commit();
      __syncthreads();
    }
  }
  double* ybase = Y + (int64_t)tl.split * split_stride + sg.yoff;
#pragma unroll
  for (int f = 0; f < RF; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tl.row0 + w * 16 * RF + f * 16 + mfma::f64_row(lane, v);
        const int col = t * 16 + mfma::f64_col(lane, v);
        if (row < sg.outlen && col < kcols) ybase[row + (int64_t)(cbase + col) * ldy] = acc[f][t][v];
      }
}

// X (FP64) -> SPLIT BF16 planes with the same layout, plane stride `ps`.
template <int SPLIT>
__global__ void k_split_bf16(const double* __restrict__ X, int64_t n, __bf16* __restrict__ Xs,
                             int64_t ps) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = (float)X[i];
    const __bf16 h = (__bf16)v;
    Xs[i] = h;
    if constexpr (SPLIT >= 2) {
      const float r1 = v - (float)h;
      const __bf16 m = (__bf16)r1;
      Xs[i + ps] = m;
      if constexpr (SPLIT >= 3) Xs[i + 2 * ps] = (__bf16)(r1 - (float)m);
    }
  }
}

// Exact split of an FP32 value into SPLIT BF16 terms by truncation:
// hi = top 8 significant bits, mid = top 8 bits of v - hi, lo = the rest,
// each term's bits being the upper half of an FP32 word.  With SPLIT = 3
// hi + mid + lo = v exactly (8 + 8 + 8 = 24 bits); no conversions needed.
template <int SPLIT>
__device__ inline void split_bf16_bits(float v, uint32_t (&o)[3]) {
  const uint32_t h = __float_as_uint(v) & 0xFFFF0000u;
  o[0] = h;
  if constexpr (SPLIT >= 2) {
    const float r1 = v - __uint_as_float(h);
    const uint32_t m = __float_as_uint(r1) & 0xFFFF0000u;
    o[1] = m;
    if constexpr (SPLIT >= 3) o[2] = __float_as_uint(r1 - __uint_as_float(m));
  }
}

// No K-tile of 64 reduction rows starting at kb meets the excluded interval.
// Rows past the segment end and reduction rows past the tile end are clamped
// to the last valid index, so every generated entry is an entry of A: the
// rows are never stored and the X rows are zero in LDS.  X rows inside the
// excluded interval are zero too (tiles that touch it are staged element by
// element); that interval is the node's own diagonal block of A, so its
// entries are finite whenever A's are.
template <class Gen, bool TRANS, int SPLIT>
__global__ __launch_bounds__(kWGs) void k_sketch_lp(Gen gen, const SegSk* __restrict__ segs,
                                                    const SkTile* __restrict__ tiles,
                                                    const __bf16* __restrict__ Xs, int64_t ldx,
                                                    int64_t ps, double* __restrict__ Y, int64_t ldy,
                                                    int64_t split_stride) {
  // rows of X padded to 160 bytes: the B-fragment reads (ds_read_b128, lane
  // l at row l % 16, byte 16 (l / 16)) hit distinct banks; with 144-byte rows
  // they conflict (4 extra cycles per read, measured)
  constexpr int RF = 2, NT = 4, NC = 64, KT = 64, KP = KT + 16;
  // the single-term pass keeps FP32 sums (its error is that of BF16 entries);
  // the three-term pass flushes them to FP64 every 8 K-tiles
  constexpr int FLUSH = SPLIT > 1 ? 8 : 0;
  // SPLIT > 1: two K-tile buffers, as in k_sketch (one barrier per tile).
  // The single-term pass keeps one buffer and two barriers per tile: the
  // second buffer's address costs it a register and a wave per SIMD.
  constexpr int NB = SPLIT > 1 ? 2 : 1;
  __shared__ __attribute__((aligned(16))) __bf16 xs[NB][SPLIT][NC][KP];
  __shared__ __attribute__((aligned(16))) float rp[NB][KT];
  const SkTile tl = tiles[blockIdx.x];
  const SegSk sg = segs[tl.seg];
  const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
  const int cbase = blockIdx.y * NC;
  const int kcols = min(NC, sg.k - cbase);
  if (kcols <= 0) return;  // a column group past this segment's width: nothing to store
  const bool aligned = ((sg.xoff & 7) == 0) && ((ldx & 7) == 0);
  mfma::v4d acc64[RF][NT];
  mfma::v4f acc32[RF][NT];
#pragma unroll
  for (int f = 0; f < RF; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      acc64[f][t] = mfma::v4d{0.0, 0.0, 0.0, 0.0};
      acc32[f][t] = mfma::v4f{0.f, 0.f, 0.f, 0.f};
    }
  int64_t iglob[RF];
  float opt[RF];
#pragma unroll
  for (int f = 0; f < RF; ++f) {
    const int rloc = tl.row0 + w * 32 + f * 16 + (lane % 16);
    iglob[f] = sg.out0 + min(rloc, max(sg.outlen - 1, 0));
    opt[f] = TRANS ? gen.col_pt(iglob[f]) : gen.row_pt(iglob[f]);
  }
  const int kq = lane / 16;
  int nt = 0;
  // Software pipeline: the next K-tile's BF16 planes and points are loaded
  // into registers while the current tile is computed from LDS.  Tiles that
  // are partial or unaligned are loaded element-wise at commit time.
  constexpr int NPRE = SPLIT * NC * (KT / 8) / kWGs;
  static_assert(SPLIT * NC * (KT / 8) % kWGs == 0, "prefetch layout");
  uint4 pre[NPRE];
  float ppt = 0.f;
  auto issue = [&](int kb) -> bool {
    const bool vec = aligned && ktile_full(sg, kb, tl.re);
    if (vec) {
#pragma unroll
      for (int i = 0; i < NPRE; ++i) {
        const int e = threadIdx.x + i * kWGs;
        const int ch = e % (KT / 8), cc = (e / (KT / 8)) % NC, sp = e / ((KT / 8) * NC);
        // in bounds for every lane (clamped column), masked by a select
        const uint4 v = *reinterpret_cast<const uint4*>(Xs + sp * ps + sg.xoff + kb + 8 * ch +
                                                        (int64_t)(cbase + min(cc, kcols - 1)) * ldx);
        pre[i] = cc < kcols ? v : make_uint4(0, 0, 0, 0);
      }
    }
    if (threadIdx.x < KT) {
      const int64_t jg = sg.red0 + min(kb + (int)threadIdx.x, tl.re - 1);
      ppt = TRANS ? gen.row_pt(jg) : gen.col_pt(jg);
    }
    return vec;
  };
  auto commit = [&](int buf, int kb, bool vec) {
    if (vec) {
#pragma unroll
      for (int i = 0; i < NPRE; ++i) {
        const int e = threadIdx.x + i * kWGs;
        const int ch = e % (KT / 8), cc = (e / (KT / 8)) % NC, sp = e / ((KT / 8) * NC);
        *reinterpret_cast<uint4*>(&xs[buf][sp][cc][8 * ch]) = pre[i];
      }
    } else {
      for (int e = threadIdx.x; e < SPLIT * NC * KT; e += kWGs) {
        const int kk = e % KT, cc = (e / KT) % NC, sp = e / (KT * NC), j = kb + kk;
        const bool excl = (uint64_t)(sg.red0 + j - sg.excl0) < (uint64_t)sg.excllen;
        xs[buf][sp][cc][kk] = (j < tl.re && cc < kcols && !excl)
                             ? Xs[sp * ps + sg.xoff + j + (int64_t)(cbase + cc) * ldx]
                             : (__bf16)0.f;
      }
    }
    if (threadIdx.x < KT) rp[buf][threadIdx.x] = ppt;
  };
  // Entries 2jp, 2jp+1 of this lane's eight in K-step ks, both row
  // fragments: word jp of every term (two entries per 32-bit word).
  auto gen_pair = [&](int buf, int kb, int ks, int jp, v4u (&wd)[RF][SPLIT]) {
    const int k0 = ks * 32 + 8 * kq;
    float pts[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) pts[h] = rp[buf][k0 + 2 * jp + h];
#pragma unroll
    for (int f = 0; f < RF; ++f) {
      float val[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int jj = 2 * jp + h;
        const int64_t jg = sg.red0 + min(kb + k0 + jj, tl.re - 1);
        val[h] = TRANS ? gen.eval(pts[h], opt[f], jg, iglob[f]) : gen.eval(opt[f], pts[h], iglob[f], jg);
      }
      if constexpr (SPLIT == 1) {
        // the leading term of the exact split (truncation): the upper
        // halves of both words in one v_perm.  Rounding to nearest has
        // the same cost and the same error on scattered points, but on
        // an equispaced Cauchy grid (a Toeplitz matrix) it made the
        // final basis 6x less accurate.
        wd[f][0][jp] = __builtin_amdgcn_perm(__float_as_uint(val[1]), __float_as_uint(val[0]), 0x07060302u);
      } else {
        uint32_t o[2][3];
        split_bf16_bits<SPLIT>(val[0], o[0]);
        split_bf16_bits<SPLIT>(val[1], o[1]);
#pragma unroll
        for (int sp = 0; sp < SPLIT; ++sp) wd[f][sp][jp] = __builtin_amdgcn_perm(o[1][sp], o[0][sp], 0x07060302u);
      }
    }
  };
  // The products of K-step ks with column tile t.
  auto mma_tile = [&](int buf, int ks, int t, const v4u (&wd)[RF][SPLIT]) {
    const int k0 = ks * 32 + 8 * kq;
    mfma::v8bf b[SPLIT];
#pragma unroll
    for (int sp = 0; sp < SPLIT; ++sp)
      b[sp] = *reinterpret_cast<const mfma::v8bf*>(&xs[buf][sp][t * 16 + (lane % 16)][k0]);
#pragma unroll
    for (int sa = 0; sa < SPLIT; ++sa)
#pragma unroll
      for (int sb = 0; sb < SPLIT - sa; ++sb)
#pragma unroll
        for (int f = 0; f < RF; ++f)
          acc32[f][t] = mfma::bf16_16x16x32(__builtin_bit_cast(mfma::v8bf, wd[f][sa]), b[sb], acc32[f][t]);
  };
  // One K-tile of two K-steps: each step's entries, then its products.
  // (Interleaving the second step's entry generation with the first step's
  // products, down to one MFMA per three vector instructions, measured no
  // faster: 39.4 -> 39.7 ms for the three-term pass.  A wave's BF16 MFMAs
  // and its vector instructions do not overlap here; the pass costs the
  // sum of the two, 768 + ~670 cycles per K-step, as its counters show.)
  auto tile = [&](int buf, int kb) {
    static_assert(KT == 64 && NT == 4, "two K-steps, four column tiles");
    v4u w[RF][SPLIT];
#pragma unroll
    for (int ks = 0; ks < KT / 32; ++ks) {
#pragma unroll
      for (int jp = 0; jp < 4; ++jp) gen_pair(buf, kb, ks, jp, w);
#pragma unroll
      for (int t = 0; t < NT; ++t) mma_tile(buf, ks, t, w);
    }
  };
  commit(0, tl.rb, issue(tl.rb));
  __syncthreads();
  int cur = 0;
  for (int kb = tl.rb; kb < tl.re; kb += KT) {
    const bool has_next = kb + KT < tl.re;
    const bool vec_next = has_next ? issue(kb + KT) : false;
    if constexpr (NB == 2) {
      tile(cur, kb);
      if (has_next) commit(cur ^ 1, kb + KT, vec_next);
      __syncthreads();
      cur ^= 1;
    } else {
      tile(0, kb);
      __syncthreads();
      if (has_next) {
        commit(0, kb + KT, vec_next);
        __syncthreads();
      }
    }
    if constexpr (FLUSH > 0) {
      if (++nt == FLUSH) {
        nt = 0;
#pragma unroll
        for (int f = 0; f < RF; ++f)
#pragma unroll
          for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              acc64[f][t][v] += (double)acc32[f][t][v];
              acc32[f][t][v] = 0.f;
            }
      }
    }
  }
  double* ybase = Y + (int64_t)tl.split * split_stride + sg.yoff;
#pragma unroll
  for (int f = 0; f < RF; ++f)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = tl.row0 + w * 32 + f * 16 + mfma::bf16_row(lane, v);
        const int col = t * 16 + mfma::bf16_col(lane, v);
        if (row < sg.outlen && col < kcols)
          ybase[row + (int64_t)(cbase + col) * ldy] = acc64[f][t][v] + (double)acc32[f][t][v];
      }
}

struct SumTile {
  int32_t seg, row0;
};
__global__ __launch_bounds__(kWGs) void k_sketch_sum(const SegSk* __restrict__ segs,
                                                     const SumTile* __restrict__ tiles,
                                                     const double* __restrict__ part, int nsplit,
                                                     int64_t split_stride, double* __restrict__ Y,
                                                     int64_t ldy) {
  const SumTile tl = tiles[blockIdx.x];
  const SegSk sg = segs[tl.seg];
  const int r = tl.row0 + (threadIdx.x % 64);
  if (r >= sg.outlen) return;
  for (int c = threadIdx.x / 64; c < sg.k; c += 4) {
    const int64_t o = sg.yoff + r + (int64_t)c * ldy;
    double s = 0.0;
    for (int p = 0; p < nsplit; ++p) s += part[(int64_t)p * split_stride + o];
    Y[o] = s;
  }
}

struct SketchPlan {
  std::vector<SkTile> tiles;
  int nsplit = 1, ncg = 1;
  int64_t span = 0;
};

static SketchPlan plan_sketch(const std::vector<SegSk>& segs, int kcap, int ncols_per_wg, int rows_per_wg,
                              int64_t ldy, int64_t want) {
  SketchPlan p;
  p.ncg = (int)ceil_div(round_up(kcap, 16), ncols_per_wg);
  int64_t nrow_tiles = 0, redmax = 0;
  for (auto& sg : segs) {
    nrow_tiles += ceil_div(std::max(sg.outlen, 0), rows_per_wg);
    redmax = std::max<int64_t>(redmax, sg.redlen);
  }
  // split the reduction when the row tiles alone cannot fill the chip
  // (`want` workgroups: 1024 by default, or the autotuner's choice)
  if (nrow_tiles * p.ncg < want) {
    p.nsplit = (int)std::min<int64_t>(ceil_div(want, std::max<int64_t>(nrow_tiles * p.ncg, 1)),
                                      std::max<int64_t>(1, redmax / 256));
    p.nsplit = std::max(1, std::min(p.nsplit, 64));
    // the partial sums span the output's extent once per split: at most
    // 2^28 doubles
    int64_t span = 0;
    for (auto& sg : segs) span = std::max<int64_t>(span, sg.yoff + sg.outlen + (int64_t)(sg.k - 1) * ldy + 1);
    p.nsplit = (int)std::max<int64_t>(1, std::min<int64_t>(p.nsplit, ((int64_t)1 << 28) / std::max<int64_t>(span, 1)));
  }
  for (size_t i = 0; i < segs.size(); ++i) {
    const auto& sg = segs[i];
    if (sg.outlen <= 0 || sg.k <= 0) continue;
    const int64_t chunk = round_up((int)ceil_div(std::max(sg.redlen, 1), p.nsplit), 64);
    for (int r0 = 0; r0 < sg.outlen; r0 += rows_per_wg)
      for (int q = 0; q < p.nsplit; ++q) {
        const int rb = (int)std::min<int64_t>((int64_t)q * chunk, sg.redlen);
        const int re = (int)std::min<int64_t>((int64_t)(q + 1) * chunk, sg.redlen);
        p.tiles.push_back({(int32_t)i, r0, rb, re, q, 0});
      }
  }
  if (p.nsplit > 1)
    for (auto& sg : segs)
      p.span = std::max<int64_t>(p.span, sg.yoff + sg.outlen + (int64_t)(sg.k - 1) * ldy + 1);
  return p;
}

// Tiles of the split-K sum (64 rows of one segment each).
static std::vector<SumTile> sum_tiles(const std::vector<SegSk>& segs) {
  std::vector<SumTile> st;
  for (size_t i = 0; i < segs.size(); ++i)
    for (int r0 = 0; r0 < segs[i].outlen; r0 += 64) st.push_back({(int32_t)i, r0});
  return st;
}

template <class Gen, bool TRANS>
static void launch_sketch(const Gen& gen, int NT, int RF, dim3 grid, const SegSk* ds, const SkTile* dt,
                          const double* X, int64_t ldx, double* Y, int64_t ldy, int64_t ss,
                          hipStream_t s) {
#define HMAT_SK(NT_, RF_) k_sketch<Gen, TRANS, NT_, RF_><<<grid, kWGs, 0, s>>>(gen, ds, dt, X, ldx, Y, ldy, ss)
  if (RF == 4) {
    switch (NT) {
      case 1: HMAT_SK(1, 4); break;
      case 2: HMAT_SK(2, 4); break;
      case 3: HMAT_SK(3, 4); break;
      default: HMAT_SK(4, 4); break;
    }
  } else {
    switch (NT) {
      case 1: HMAT_SK(1, 2); break;
      case 2: HMAT_SK(2, 2); break;
      case 3: HMAT_SK(3, 2); break;
      default: HMAT_SK(4, 2); break;
    }
  }
#undef HMAT_SK
}

// One sketch with a given tiling: rows per workgroup 64 * RF, split-K target.
static void run_sketch(const GenDesc& g, bool trans, const std::vector<SegSk>& segs, const double* X,
                       int64_t ldx, double* Y, int64_t ldy, int kcap, Workspace& ws, int RF, int64_t want,
                       bool prof_on, hipStream_t s) {
  const int kpad = round_up(kcap, 16);
  const int NT = std::min(4, kpad / 16);
  const int rows_per_wg = 64 * RF;
  SketchPlan p = plan_sketch(segs, kcap, NT * 16, rows_per_wg, ldy, want);
  if (p.tiles.empty()) return;
  // segments, tiles and (split-K) sum tiles staged together
  const std::vector<SumTile> st = p.nsplit > 1 ? sum_tiles(segs) : std::vector<SumTile>();
  StageGroup sgp;
  const int is = sgp.add(segs), it = sgp.add(p.tiles), isum = sgp.add(st);
  sgp.put(s);
  const SegSk* ds = sgp.get<SegSk>(is);
  const SkTile* dt = sgp.get<SkTile>(it);
  double* out = Y;
  if (p.nsplit > 1) {
    ws.part.reserve((size_t)(p.span * p.nsplit));
    out = ws.part.get();
  }
  dim3 grid((unsigned)p.tiles.size(), (unsigned)p.ncg);
  double flops = 0.0;
  for (auto& sg : segs) flops += 2.0 * sg.outlen * (double)sg.redlen * sg.k;
  ProfScope prof(g.kind == GenKind::Dense ? "sketch_dense" : "sketch_gen", s, prof_on ? flops : 0.0);
  if (!prof_on) prof.cancel();
  with_gen(g, [&](auto gen) {
    if (trans)
      launch_sketch<decltype(gen), true>(gen, NT, RF, grid, ds, dt, X, ldx, out, ldy, p.span, s);
    else
      launch_sketch<decltype(gen), false>(gen, NT, RF, grid, ds, dt, X, ldx, out, ldy, p.span, s);
  });
  HIP_CHECK(hipGetLastError());
  prof.finish();
  if (p.nsplit > 1) {
    k_sketch_sum<<<(unsigned)st.size(), kWGs, 0, s>>>(ds, sgp.get<SumTile>(isum), ws.part.get(), p.nsplit, p.span,
                                                     Y, ldy);
    HIP_CHECK(hipGetLastError());
  }
}

// Shape class of a sketch launch for the autotuner.
static TuneKey sketch_key(int kind, bool trans, const std::vector<SegSk>& segs, int kcap) {
  int64_t rows = 0, red = 0;
  for (auto& sg : segs) {
    rows += sg.outlen;
    red += sg.redlen;
  }
  auto lg = [](int64_t v) { return v > 0 ? 64 - __builtin_clzll((unsigned long long)v) : 0; };
  return TuneKey{kind, trans ? 1 : 0, round_up(kcap, 16), lg(rows), lg(red / std::max<size_t>(segs.size(), 1)),
                 lg((int64_t)segs.size())};
}

// Time `ncand` variants of one launch on the stream and return the fastest
// (the last variant run leaves its result in place; the caller re-runs the
// winner when it was not the last).
template <class Run>
static int time_candidates(int ncand, Run&& run, hipStream_t s) {
  hipEvent_t a, b;
  HIP_CHECK(hipEventCreate(&a));
  HIP_CHECK(hipEventCreate(&b));
  int best = 0;
  float best_ms = 0.f;
  for (int c = 0; c < ncand; ++c) {
    HIP_CHECK(hipEventRecord(a, s));
    run(c);
    HIP_CHECK(hipEventRecord(b, s));
    HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    if (c == 0 || ms < best_ms) {
      best = c;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

void sketch(const GenDesc& g, bool trans, const std::vector<SegSk>& segs, const double* X, int64_t ldx,
            double* Y, int64_t ldy, int kcap, Workspace& ws, hipStream_t s) {
  if (segs.empty() || kcap <= 0) return;
  int64_t rows = 0;
  for (auto& sg : segs) rows += sg.outlen;
  // default: 256-row tiles when the segments are long enough to fill the
  // chip with them, split-K up to 1024 workgroups
  int RF = (rows / 256 >= 512) ? 4 : 2;
  int64_t want = 1024;
  if (autotune_enabled()) {
    static const int kRF[2] = {2, 4};
    static const int64_t kWant[3] = {512, 1024, 2048};
    const TuneKey key = sketch_key(g.kind == GenKind::Dense ? 0 : 1, trans, segs, kcap);
    int choice = -1;
    if (!tune_lookup(key, choice)) {
      choice = time_candidates(6, [&](int c) {
        run_sketch(g, trans, segs, X, ldx, Y, ldy, kcap, ws, kRF[c / 3], kWant[c % 3], false, s);
      }, s);
      tune_store(key, choice);
      if (choice == 5) return;  // the last candidate's result is in Y
    }
    RF = kRF[choice / 3];
    want = kWant[choice % 3];
  }
  run_sketch(g, trans, segs, X, ldx, Y, ldy, kcap, ws, RF, want, true, s);
}

template <class Gen, bool TRANS>
static void launch_sketch_lp(const Gen& gen, int split, dim3 grid, const SegSk* ds, const SkTile* dt,
                             const __bf16* Xs, int64_t ldx, int64_t ps, double* Y, int64_t ldy,
                             int64_t ss, hipStream_t s) {
  switch (split) {
    case 1: k_sketch_lp<Gen, TRANS, 1><<<grid, kWGs, 0, s>>>(gen, ds, dt, Xs, ldx, ps, Y, ldy, ss); break;
    case 2: k_sketch_lp<Gen, TRANS, 2><<<grid, kWGs, 0, s>>>(gen, ds, dt, Xs, ldx, ps, Y, ldy, ss); break;
    default: k_sketch_lp<Gen, TRANS, 3><<<grid, kWGs, 0, s>>>(gen, ds, dt, Xs, ldx, ps, Y, ldy, ss); break;
  }
}

void sketch_lp(const GenDesc& g, bool trans, int split, const std::vector<SegSk>& segs,
               const double* X, int64_t ldx, double* Y, int64_t ldy, int kcap,
               Workspace& ws, hipStream_t s) {
  if (segs.empty() || kcap <= 0) return;
  split = std::max(1, std::min(split, 3));
  // BF16 planes of X over the rows any segment reduces over
  int64_t xend = 0;
  for (auto& sg : segs) xend = std::max<int64_t>(xend, sg.xoff + sg.redlen + (int64_t)(kcap - 1) * ldx);
  const int64_t ps = round_up((int)std::min<int64_t>(xend, (int64_t)1 << 30), 64);
  HMAT_REQUIRE(xend < ((int64_t)1 << 30), HMAT_ERR_ARG, "sketch operand too large for BF16 planes");
  ws.bf.reserve((size_t)(ps * split));
  {
    const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(xend, kWGs), 8192);
    switch (split) {
      case 1: k_split_bf16<1><<<gx, kWGs, 0, s>>>(X, xend, ws.bf.get(), ps); break;
      case 2: k_split_bf16<2><<<gx, kWGs, 0, s>>>(X, xend, ws.bf.get(), ps); break;
      default: k_split_bf16<3><<<gx, kWGs, 0, s>>>(X, xend, ws.bf.get(), ps); break;
    }
    HIP_CHECK(hipGetLastError());
  }
  auto run = [&](int64_t want, bool prof_on) {
    SketchPlan p = plan_sketch(segs, kcap, 64, 128, ldy, want);
    if (p.tiles.empty()) return;
    const std::vector<SumTile> st = p.nsplit > 1 ? sum_tiles(segs) : std::vector<SumTile>();
    StageGroup sgp;
    const int is = sgp.add(segs), it = sgp.add(p.tiles), isum = sgp.add(st);
    sgp.put(s);
    const SegSk* ds = sgp.get<SegSk>(is);
    const SkTile* dt = sgp.get<SkTile>(it);
    double* out = Y;
    if (p.nsplit > 1) {
      ws.part.reserve((size_t)(p.span * p.nsplit));
      out = ws.part.get();
    }
    dim3 grid((unsigned)p.tiles.size(), (unsigned)p.ncg);
    double flops = 0.0;
    for (auto& sg : segs) flops += 2.0 * sg.outlen * (double)sg.redlen * sg.k;
    static const char* const kProfName[3] = {"sketch_lp1", "sketch_lp2", "sketch_lp"};
    ProfScope prof(kProfName[split - 1], s, flops);
    if (!prof_on) prof.cancel();
    with_gen_lp(g, [&](auto gen) {
      if (trans)
        launch_sketch_lp<decltype(gen), true>(gen, split, grid, ds, dt, ws.bf.get(), ldx, ps, out,
                                              ldy, p.span, s);
      else
        launch_sketch_lp<decltype(gen), false>(gen, split, grid, ds, dt, ws.bf.get(), ldx, ps,
                                               out, ldy, p.span, s);
    });
    HIP_CHECK(hipGetLastError());
    prof.finish();
    if (p.nsplit > 1) {
      k_sketch_sum<<<(unsigned)st.size(), kWGs, 0, s>>>(ds, sgp.get<SumTile>(isum), ws.part.get(), p.nsplit,
                                                       p.span, Y, ldy);
      HIP_CHECK(hipGetLastError());
    }
  };
  int64_t want = 1024;
  if (autotune_enabled()) {
    static const int64_t kWant[3] = {512, 1024, 2048};
    const TuneKey key = sketch_key(2 + split, trans, segs, kcap);
    int choice = -1;
    if (!tune_lookup(key, choice)) {
      choice = time_candidates(3, [&](int c) { run(kWant[c], false); }, s);
      tune_store(key, choice);
      if (choice == 2) return;  // the last candidate's result is in Y
    }
    want = kWant[choice];
  }
  run(want, true);
}

}  // namespace kern
}  // namespace hmat
