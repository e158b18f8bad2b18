// Launch wrappers for the gfx950 kernels of the HODLR / HSS engine.
//
// Storage model: every tall operand is a column-major device array with a
// leading dimension (usually N); a "segment" is a contiguous run of rows of
// such an array that belongs to one tree node.  All batched kernels take a
// device array of per-segment descriptors so that one launch covers every
// node of a tree level.
#pragma once
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace hmat {

// Y(seg rows, 0:q) = alpha * X(seg rows, 0:p) * T_s(0:p, 0:q) + beta * Y
struct SegNN {
  int64_t xoff, toff, yoff;
  int32_t len, p, q, pad_;
};

// O_s(0:p, 0:q) = alpha * X(seg rows, 0:p)^T * Y(seg rows, 0:q) + beta * O_s
struct SegTN {
  int64_t xoff, yoff, ooff;
  int32_t len, p, q, pad_;
};

// Sketch of one off-diagonal block A(rows, cols) against a tall operand:
//   trans = false:  Y(out rows) = A(out rows, red) * X(red rows)
//   trans = true:   Y(out rows) = A(red, out rows)^T * X(red rows)
// out/red are global index ranges; entries whose reduction index falls in
// [excl0, excl0 + excllen) count as zero (HSS row blocks).
struct SegSk {
  int64_t out0, red0, excl0, xoff, yoff;
  int32_t outlen, redlen, excllen, k;
};

// Small square matrices, one per batch entry, stored column-major at
// base + off with leading dimension ld.
struct SmallMat {
  int64_t off;
  int32_t n, ld;
};

enum class GenKind : int {
  Dense = 0,
  Gaussian = 1,     // exp(-(x_i-x_j)^2/(2 p0^2)) + p1 delta
  Cauchy = 2,       // 1/(x_i - y_j)
  RefCauchy = 3,    // 1/(x_i + x_j + 1) + p1 delta
  RefGaussian = 4,  // exp(-(x_i-x_j)^2/p0^2) + p1 delta
  RefCoulomb = 5,   // 1/|x_i-x_j|, diag 0, + p1 delta
  GaussianDSigma = 6  // d/dsigma of Gaussian: exp(-(x_i-x_j)^2/(2 p0^2)) (x_i-x_j)^2 / p0^3
};

struct GenDesc {
  GenKind kind = GenKind::Dense;
  const double* A = nullptr;  // dense
  int64_t lda = 0;
  const double* x = nullptr;  // points
  const double* y = nullptr;
  double p0 = 0, p1 = 0;
};

// Dense blocks out of a generator: out[off + i + j*ld] = A(r0+i, c0+j)
struct BlockDesc {
  int64_t r0, c0, off;
  int32_t m, n, ld, pad_;
};

namespace kern {

// scratch used by two-stage reductions
struct Workspace {
  DevBuf<double> part;
  DevBuf<float> part32;
  DevBuf<__bf16> bf;  // BF16 planes of the mixed-precision sketch operand
  DevBuf<int64_t> aux;
  template <class T> DevBuf<T>& parts();
  void addresses(std::vector<const void*>& v) const {
    v.insert(v.end(), {part.get(), part32.get(), bf.get(), aux.get()});
  }
};
template <> inline DevBuf<double>& Workspace::parts<double>() { return part; }
template <> inline DevBuf<float>& Workspace::parts<float>() { return part32; }

struct NNTile {
  int32_t seg, row0, c0, pad_;
};
struct TNItem {
  int32_t seg, r0, r1, pt, qt, slot;
};
struct TNOut {
  int32_t seg, pt, qt, slot0, nslot, pad_;
};

// Launch plans: descriptors built and uploaded once, launched any number of
// times with no allocation, copy or host synchronisation (graph capture).
struct NNPlan {
  DevBuf<SegNN> segs;
  DevBuf<NNTile> tiles;
  int ntiles = 0, mode = 0;
};
struct TNPlan {
  DevBuf<SegTN> segs;
  DevBuf<TNItem> items;
  DevBuf<TNOut> outs;
  int nitems = 0, nouts = 0, nslots = 0;  // launch_tn needs nslots * 4096 partials
};
// Tilings and launches on descriptors already on the device (callers that
// upload several launches' descriptors together).  nn_tiles returns the
// kernel variant, trsm_tiles the largest k, tn_items the partial slots.
int nn_tiles(const std::vector<SegNN>& segs, std::vector<NNTile>& tiles);
template <class F>
void launch_nn_raw(int mode, int ntiles, const SegNN* ds, const NNTile* dt, const F* X, int64_t ldx, const F* T,
                   int64_t ldt, F* Y, int64_t ldy, F alpha, F beta, hipStream_t s);
int trsm_tiles(const std::vector<SegNN>& segs, std::vector<NNTile>& tiles);
void launch_trsm_raw(int kmax, int ntiles, const SegNN* ds, const NNTile* dt, const double* X, int64_t ldx,
                     const double* R, int64_t ldr, double* Y, int64_t ldy, hipStream_t s);
int tn_items(const std::vector<SegTN>& segs, std::vector<TNItem>& items, std::vector<TNOut>& outs);
template <class F>
void launch_tn_raw(int nitems, int nouts, const SegTN* ds, const TNItem* di, const TNOut* dout, const F* X,
                   int64_t ldx, const F* Y, int64_t ldy, F* O, int64_t ldo, F alpha, F beta, F* part,
                   hipStream_t s);

void plan_nn(const std::vector<SegNN>& segs, NNPlan& p, hipStream_t s);
template <class F>
void launch_nn(const NNPlan& p, const F* X, int64_t ldx, const F* T, int64_t ldt, F* Y, int64_t ldy, F alpha,
               F beta, hipStream_t s);
void plan_tn(const std::vector<SegTN>& segs, TNPlan& p, hipStream_t s);
template <class F>
void launch_tn(const TNPlan& p, const F* X, int64_t ldx, const F* Y, int64_t ldy, F* O, int64_t ldo, F alpha,
               F beta, F* part, hipStream_t s);

// F = double or float (FP32 factorisation of the mixed-precision path)
template <class F>
void seg_gemm_nn(const std::vector<SegNN>& segs, const F* X, int64_t ldx,
                 const F* T, int64_t ldt, F* Y, int64_t ldy, F alpha, F beta, hipStream_t s);

// Y(seg rows, 0:q) = X(seg rows, 0:q) R_s^{-1} for upper-triangular R_s
// (q x q at toff, leading dimension ldt): row-wise forward substitution,
// backward stable (unlike multiplying by an explicit inverse, whose error
// grows with cond(R)).  Uses xoff/toff/yoff/len/q of each SegNN.
void seg_trsm_upper(const std::vector<SegNN>& segs, const double* X,
                    int64_t ldx, const double* R, int64_t ldr, double* Y, int64_t ldy, hipStream_t s);
// The same with its descriptors uploaded once and launched any number of
// times.  For k <= 64 each row is read completely before it is written, so
// X == Y (in place) is allowed there.
struct TrsmPlan {
  DevBuf<SegNN> segs;
  DevBuf<NNTile> tiles;
  int ntiles = 0, kmax = 0;
};
void plan_trsm(const std::vector<SegNN>& segs, TrsmPlan& p, hipStream_t s);
void launch_trsm(const TrsmPlan& p, const double* X, int64_t ldx, const double* R, int64_t ldr, double* Y,
                 int64_t ldy, hipStream_t s);

template <class F>
void seg_gemm_tn(const std::vector<SegTN>& segs, const F* X, int64_t ldx,
                 const F* Y, int64_t ldy, F* O, int64_t ldo, F alpha, F beta, Workspace& ws,
                 hipStream_t s);

void sketch(const GenDesc& g, bool trans, const std::vector<SegSk>& segs, const double* X, int64_t ldx,
            double* Y, int64_t ldy, int kcap, Workspace& ws, hipStream_t s);

// Same contraction on BF16 MFMA: entries generated in FP32 and split into
// `split` (1..3) BF16 terms, the tall operand likewise; products of terms
// whose split indices sum to < split are accumulated in FP32 and flushed to
// FP64 every 512 reduction indices.  Y is FP64.
void sketch_lp(const GenDesc& g, bool trans, int split, const std::vector<SegSk>& segs,
               const double* X, int64_t ldx, double* Y, int64_t ldy, int kcap,
               Workspace& ws, hipStream_t s);

// Gram G (n x n per entry) -> upper R with R^T R = G + shift_s I,
// shift_s = shift_coef * trace(G) when shift_coef > 0.  An entry with
// shift_coef = 0 whose factorisation breaks down is factorised again with
// shift fallback_coef * trace(G) when fallback coefficients are given (one
// per entry).  info[b] != 0 when the last attempt broke down; gerr[b] =
// max |G - I| (0 for an all-zero G), the orthogonality defect of the
// columns the Gram came from.
void chol_shifted(const std::vector<SmallMat>& mats, double* G, double* R, const std::vector<double>& shift_coef,
                  int* info, double* gerr, hipStream_t s,
                  const std::vector<double>& fallback_coef = std::vector<double>());
// The same on descriptors and coefficients already on the device
// (dfallback may be null).
void launch_chol_shifted(const SmallMat* dmats, int nmats, int nmax, double* G, double* R, const double* dshift,
                         int* info, double* gerr, hipStream_t s, const double* dfallback = nullptr);

// One-sided Jacobi SVD of n x n matrices A = U diag(S) V^T (A overwritten
// by U).  S sorted descending, U/V columns permuted accordingly.  A pair of
// columns is rotated while |a_p . a_q| > 1e-15 |a_p| |a_q| and
// |a_p . a_q| > floor_rel ||A||_F^2: the default floor only guards against
// underflow; a larger one stops resolving columns whose norms are below
// sqrt(floor_rel) ||A||_F against each other (singular values that the
// input does not determine to better than that anyway).  A V = U diag(S)
// holds in either case.
void jacobi_svd(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, double* A, double* V,
                double* S, int64_t sstride, double* work, hipStream_t s, double floor_rel = 1e-30);

// In-place inverse by Gauss-Jordan with partial pivoting.
template <class F>
void gj_inverse(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, F* A, int* info,
                hipStream_t s);

// Y(rows, c) = 0 for all (rows,c) of the rectangle; identity helpers.
template <class F>
void fill_identity_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, F* Y, int64_t ldy,
                        hipStream_t s);

// Device Omega: Y[yoff + r + c*ldy] for r < len, c < k from a counter-based
// generator keyed by (seed, stream id, r, c); uniform [0,1) or N(0,1).
void fill_random_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, double* Y,
                      int64_t ldy, uint64_t seed, int dist, hipStream_t s);

void gen_blocks(const GenDesc& g, const std::vector<BlockDesc>& blocks, DevBuf<BlockDesc>& db,
                double* out, hipStream_t s);

// In-place transpose of block i (lens[i] x lens[i] at base + i*ld*ld, ld ld).
void transpose_blocks(double* base, int64_t ld, const std::vector<int32_t>& lens, hipStream_t s);

// Y(seg rows, 0:q) = X(seg rows, 0:q) (xoff/yoff/len/q of each SegNN)
void copy_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, const double* X, int64_t ldx,
               double* Y, int64_t ldy, hipStream_t s);

// Truncation of a small SVD R = Ur diag(S) Vr^T to rank[b] columns, zero
// columns up to rmax:
//   s_on_ur = 0:  TU = Vr(:, :r) diag(S(:r)),  TV = Ur(:, :r)
//   s_on_ur = 1:  TU = Vr(:, :r),              TV = Ur(:, :r) diag(S(:r))
void trunc_prep(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, const double* Ur,
                const double* Vr, const double* S, int64_t sstride, const int* rank, int rmax,
                double* TU, double* TV, int s_on_ur, hipStream_t s);

// dst(0:m, 0:n) = alpha src(0:m, 0:n) + beta dst (src may be null: beta dst),
// and dst(i, i) += alpha for the identity items; per-item leading dimensions.
struct Rect {
  const double* src;
  double* dst;
  int64_t lds, ldd;
  int32_t m, n;
  double alpha, beta;
};
void rect_axpby(const std::vector<Rect>& items, DevBuf<Rect>& d, hipStream_t s);
void rect_add_identity(const std::vector<Rect>& items, DevBuf<Rect>& d, hipStream_t s);
// dst(j, i) = alpha src(i, j) + beta dst(j, i) for i < m, j < n (src m x n)
void rect_transpose(const std::vector<Rect>& items, DevBuf<Rect>& d, hipStream_t s);

// Z(:, c) = alpha X(:, c) + beta Y(:, c) for n rows, ncols columns.
void axpby_cols(int64_t n, int ncols, double alpha, const double* X, int64_t ldx, double beta,
                const double* Y, int64_t ldy, double* Z, int64_t ldz, hipStream_t s);

// Batched small GEMM C = alpha op(A) op(B) + beta C, one workgroup per
// entry (per-node work of the HSS sweeps and factorisation; sizes up to a
// few hundred).
struct BGemm {
  const double* A;
  const double* B;
  double* C;
  int32_t m, n, k, lda, ldb, ldc, ta, tb;
  double alpha, beta;
};
void bgemm(const std::vector<BGemm>& items, DevBuf<BGemm>& d, hipStream_t s);
// Node tasks: one workgroup per task runs ops[op0 .. op0 + nops) in order,
// each op reading what the previous ones wrote.  An op is a BGemm; k == 0
// makes it a copy C = alpha A + beta C (A m x n, leading dimension lda).
struct NodeTask {
  int32_t op0, nops;
};
void node_tasks(const NodeTask* dtasks, int ntasks, const BGemm* dops, hipStream_t s);
// Zero rows [r0, r0 + len) of columns [c0, c0 + ncols) of a column-major array
struct ZeroRect {
  double* p;
  int64_t ld;
  int32_t r0, len, c0, ncols;
};
void zero_rects(const std::vector<ZeroRect>& items, DevBuf<ZeroRect>& d, hipStream_t s);
// D_i(j, j) += *dshift (a device scalar) for j < dlens[i] over nleaf blocks
// of lm x lm
void add_diag(double* D, int64_t lm, const int* dlens, int nleaf, const double* dshift, hipStream_t s);
// *p = v on the stream (a kernel: recordable into a graph, value by argument)
void set_scalar(double* p, double v, hipStream_t s);

// Y = (To) X elementwise for n contiguous values.
template <class To, class Ti>
void convert(const Ti* X, To* Y, int64_t n, hipStream_t s);
// Y(:, c) = (To) X(:, c), n rows, ncols columns, explicit leading dimensions.
template <class To, class Ti>
void convert_cols(int64_t n, int ncols, const Ti* X, int64_t ldx, To* Y, int64_t ldy, hipStream_t s);
// X(:, c) += (double) D(:, c)
void add_f32_cols(int64_t n, int ncols, const float* D, int64_t ldd, double* X, int64_t ldx,
                  hipStream_t s);

// Per-column 2-norms squared of an n-row, ncols column-major array.
// part: scratch of 64 * ncols doubles
void col_norms2(const double* X, int64_t ldx, int64_t n, int ncols, double* out, double* part, hipStream_t s);

}  // namespace kern
}  // namespace hmat
