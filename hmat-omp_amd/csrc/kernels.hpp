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
  RefCoulomb = 5    // 1/|x_i-x_j|, diag 0, + p1 delta
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
  DevBuf<int64_t> aux;
};

void seg_gemm_nn(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, const double* X,
                 int64_t ldx, const double* T, int64_t ldt, double* Y, int64_t ldy,
                 double alpha, double beta, hipStream_t s);

void seg_gemm_tn(const std::vector<SegTN>& segs, DevBuf<SegTN>& dsegs, const double* X,
                 int64_t ldx, const double* Y, int64_t ldy, double* O, int64_t ldo,
                 double alpha, double beta, Workspace& ws, hipStream_t s);

void sketch(const GenDesc& g, bool trans, const std::vector<SegSk>& segs, DevBuf<SegSk>& dsegs,
            const double* X, int64_t ldx, double* Y, int64_t ldy, int kcap, Workspace& ws,
            hipStream_t s);

// Gram G (n x n per entry) -> upper R with R^T R = G + shift_s I and
// Rinv = R^{-1}.  shift_s = shift_coef * trace(G) when shift_coef > 0.
// info[b] != 0 on breakdown.
void chol_rinv(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, double* G, double* R,
               double* Rinv, const std::vector<double>& shift_coef, DevBuf<double>& dshift,
               int* info, hipStream_t s);

// One-sided Jacobi SVD of n x n matrices A = U diag(S) V^T (A overwritten
// by U).  S sorted descending, U/V columns permuted accordingly.
void jacobi_svd(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, double* A, double* V,
                double* S, int64_t sstride, double* work, hipStream_t s);

// In-place inverse by Gauss-Jordan with partial pivoting.
void gj_inverse(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, double* A,
                int* info, hipStream_t s);

// Y(rows, c) = 0 for all (rows,c) of the rectangle; identity helpers.
void fill_identity_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, double* Y,
                        int64_t ldy, hipStream_t s);

// Device Omega: Y[yoff + r + c*ldy] for r < len, c < k from a counter-based
// generator keyed by (seed, stream id, r, c); uniform [0,1) or N(0,1).
void fill_random_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, double* Y,
                      int64_t ldy, uint64_t seed, int dist, hipStream_t s);

void gen_blocks(const GenDesc& g, const std::vector<BlockDesc>& blocks, DevBuf<BlockDesc>& db,
                double* out, hipStream_t s);

// Y(seg rows, 0:q) = X(seg rows, 0:q) (xoff/yoff/len/q of each SegNN)
void copy_segs(const std::vector<SegNN>& segs, DevBuf<SegNN>& dsegs, const double* X, int64_t ldx,
               double* Y, int64_t ldy, hipStream_t s);

// Truncation of a small SVD R = Ur diag(S) Vr^T to rank[b] columns:
//   TU = Vr(:, :r) diag(S(:r)),  TV = Ur(:, :r), zero columns up to rmax.
void trunc_prep(const std::vector<SmallMat>& mats, DevBuf<SmallMat>& dm, const double* Ur,
                const double* Vr, const double* S, int64_t sstride, const int* rank, int rmax,
                double* TU, double* TV, hipStream_t s);

// Per-column 2-norms squared of an n-row, ncols column-major array.
void col_norms2(const double* X, int64_t ldx, int64_t n, int ncols, double* out, hipStream_t s);

}  // namespace kern
}  // namespace hmat
