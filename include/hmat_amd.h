/*
 * hmat_amd.h — C ABI of the MI355X (gfx950) HODLR / HSS engine.
 *
 * Drop-in boundary for the OpenMP hot path of chenxinye/hmat-omp:
 *   HODLRMatrixOMP  (include/hodlr_routines_omp.h:9-241)
 *   HSSMatrixOMP    (include/hss_routines_omp.h:11-375)
 * The reference is a header-only C++ library with no FFI layer; these entry
 * points are what a C/C++/ctypes caller binds in place of the class methods.
 * Each declaration names the method it replaces.
 *
 * Conventions
 *   - Matrices are column-major doubles with an explicit leading dimension,
 *     like the reference Matrix (include/matrix.h:36-37: data_[j*rows_+i]).
 *   - Host-pointer entry points copy in/out; *_device entry points take
 *     device pointers already resident in HBM and run on the handle's stream.
 *   - Every call returns an hmat_status; hmat_last_error() gives a message
 *     for the calling thread.  The reference ignores LAPACK info codes
 *     (include/matrix.h:129-130,139-147); this ABI reports them instead.
 *   - Calls on one handle are serialised by the caller; distinct handles
 *     are independent.
 */
#ifndef HMAT_AMD_H
#define HMAT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HMAT_OK = 0,
  HMAT_ERR_ARG = 1,         /* bad dimension / pointer / option            */
  HMAT_ERR_DEVICE = 2,      /* HIP runtime error, or no HIP device         */
  HMAT_ERR_OOM = 3,         /* device allocation failed                    */
  HMAT_ERR_SINGULAR = 4,    /* zero pivot in a leaf or capacitance block   */
  HMAT_ERR_NONFINITE = 5,   /* NaN/Inf in input or in a factor             */
  HMAT_ERR_STATE = 6,       /* call out of order (e.g. solve before build) */
  HMAT_ERR_NOT_CONVERGED = 7/* iterative refinement missed its target     */
} hmat_status;

typedef enum {
  HMAT_PREC_FP64 = 0,       /* plain FP64 everywhere                       */
  HMAT_PREC_MIXED = 1       /* low-precision sketch + FP32 factorisation,
                               FP64 projection, residual and refinement    */
} hmat_precision;

typedef enum {
  HMAT_OMEGA_UNIFORM = 0,   /* U[0,1), as Matrix::Random (matrix.h:50-57)  */
  HMAT_OMEGA_GAUSSIAN = 1   /* N(0,1)                                      */
} hmat_omega_dist;

/* HSS build modes (the `nested` argument of hmat_hss_build_*) */
typedef enum {
  HMAT_HSS_REFERENCE = 0,   /* identity transfers (hss_routines_omp.h:110-119) */
  HMAT_HSS_NESTED = 1,      /* compressed transfers, N-wide leaf blocks         */
  HMAT_HSS_SAMPLED = 2      /* compressed transfers from one sample pass        */
} hmat_hss_mode;

/* Compression options.  hmat_options_reference() gives the reference's
 * behaviour: tolerance truncation S_i > tol*S_0 (compression.h:42-46),
 * rsvd_rank 64 (hodlr_routines_omp.h:17), k = rank+10 (compression.h:74),
 * no power iterations, exact SVD when a block side is < 256
 * (compression.h:72), Omega ~ U[0,1). */
typedef struct {
  double tol;
  int32_t rank;           /* rsvd_rank (tolerance mode) or the fixed rank */
  int32_t fixed_rank;     /* 0: tolerance rule, 1: keep exactly `rank`    */
  int32_t oversample;     /* p in k = rank + p                            */
  int32_t power_iters;    /* q                                            */
  int32_t exact_below;    /* exact SVD when a block side is below this    */
  int32_t precision;      /* hmat_precision                               */
  int32_t omega_dist;     /* hmat_omega_dist (device-generated Omega)     */
  uint64_t seed;          /* device Omega seed                            */
} hmat_options;

/* Kernel matrices evaluated on the device (no dense A).  Entry (i,j):
 *   HMAT_KERNEL_GAUSSIAN     exp(-(x_i-x_j)^2 / (2 p0^2)) + p1*delta_ij
 *   HMAT_KERNEL_CAUCHY       1 / (x_i - y_j)
 *   HMAT_KERNEL_REF_CAUCHY   1 / (x_i + x_j + 1) + p1*delta_ij   (kernel.h:30-48)
 *   HMAT_KERNEL_REF_GAUSSIAN exp(-(x_i-x_j)^2 / p0^2) + p1*delta_ij (kernel.h:52-65)
 *   HMAT_KERNEL_REF_COULOMB  1/|x_i-x_j| (0 on the diagonal) + p1*delta_ij
 *                                                               (kernel.h:69-89)
 *   HMAT_KERNEL_GAUSSIAN_DSIGMA  d/dp0 of HMAT_KERNEL_GAUSSIAN (no delta term):
 *                            exp(-(x_i-x_j)^2/(2 p0^2)) (x_i-x_j)^2 / p0^3
 *                            (hmat_kernel_matvec_device only; for adjoints)
 * x (and y for CAUCHY) are host arrays of N doubles. */
typedef enum {
  HMAT_KERNEL_GAUSSIAN = 1,
  HMAT_KERNEL_CAUCHY = 2,
  HMAT_KERNEL_REF_CAUCHY = 3,
  HMAT_KERNEL_REF_GAUSSIAN = 4,
  HMAT_KERNEL_REF_COULOMB = 5,
  HMAT_KERNEL_GAUSSIAN_DSIGMA = 6
} hmat_kernel_type;

typedef struct {
  int32_t type;
  double p0;
  double p1;
  const double* x;
  const double* y;
} hmat_kernel;

typedef struct hmat_hodlr_s* hmat_hodlr;
typedef struct hmat_hss_s* hmat_hss;

/* ------------------------------------------------------------- library -- */
const char* hmat_last_error(void);
const char* hmat_version(void);
int32_t hmat_device_count(void);
void hmat_options_reference(hmat_options* opts, double tol);

/* Opt-in kernel timing (no reference counterpart; the reference has only a
 * std::chrono Timer in its drivers, src/main_hodlr.cpp:14-24).  When enabled,
 * named launches are bracketed by HIP events on their own stream and their
 * algorithmic work is summed:
 *   "sketch_gen" / "sketch_dense"  FP64 flops of the block sketch GEMMs;
 *   "matvec"                       bytes of the fused HODLR matvec;
 * and counters are summed into `work`:
 *   "chol_fallbacks" / "chol_fallback_mats"  Cholesky factorisations that
 *       broke down unshifted and were redone shifted in the same launch,
 *       over those that could;
 *   "jacobi_sweeps" / "jacobi_mats"  Jacobi SVD sweeps over matrices.
 * "orth_unconverged" is counted whether or not timing is on: calls of the
 * CholeskyQR orthogonaliser that left a segment short of Q^T Q = I after
 * its last pass (a diagnostic; the basis is kept). */
void hmat_profile_enable(int32_t on);
void hmat_profile_reset(void);
/* Tile autotuner (no reference counterpart): when on, the first sketch
 * launch of each shape class times its candidate tilings (rows per
 * workgroup, split-K target) and the fastest is kept for the process.
 * Off by default; tuned launches may round differently from the default
 * tiling.  entries() = number of shape classes decided so far. */
void hmat_autotune_enable(int32_t on);
int64_t hmat_autotune_entries(void);
/* returns 1 and fills the totals when `name` was recorded, else 0 */
int32_t hmat_profile_query(const char* name, int64_t* launches, double* ms, double* work);

/* Exact FP64 Y = A X for a generated kernel matrix, A never stored (checks
 * the true-matrix residual at sizes where a dense A cannot exist; the
 * reference forms A densely, kernel.h:30-106).  Device pointers. */
hmat_status hmat_kernel_matvec_device(const hmat_kernel* k_device_points, int64_t N,
                                      const double* dX, int64_t ldx, int32_t nrhs, double* dY,
                                      int64_t ldy);

/* ----------------------------------------------------------------- HODLR -- */
/* HODLRMatrixOMP::HODLRMatrixOMP(double tol)  hodlr_routines_omp.h:17 */
hmat_status hmat_hodlr_create(hmat_hodlr* h, const hmat_options* opts);
/* HODLRMatrixOMP::~HODLRMatrixOMP()           hodlr_routines_omp.h:18 */
hmat_status hmat_hodlr_destroy(hmat_hodlr h);

/* HODLRMatrixOMP::build_from_dense(const Matrix& A, int leaf_size)
 *   hodlr_routines_omp.h:21-79.  A: N x N host, column-major, lda >= N.
 *   omega: optional host Omega blocks in the serial draw order (post-order,
 *   A12 then A21, only blocks taking the randomised branch, each n x k
 *   column-major; hodlr_routines.h:43-53), or NULL for device Omega.
 *   omega_len must equal hmat_hodlr_omega_length() when omega != NULL. */
hmat_status hmat_hodlr_build_dense(hmat_hodlr h, const double* A, int64_t lda,
                                   int64_t N, int32_t leaf_size,
                                   const double* omega, int64_t omega_len);
/* Same, A already on the device. */
hmat_status hmat_hodlr_build_dense_device(hmat_hodlr h, const double* dA,
                                          int64_t lda, int64_t N,
                                          int32_t leaf_size,
                                          const double* omega,
                                          int64_t omega_len);
/* Matrix-free build: entries generated on the device (kernel.h:18-106 and
 * the point kernels above).  No reference counterpart for N >= 46341
 * (the reference overflows int at matrix.h:34). */
hmat_status hmat_hodlr_build_kernel(hmat_hodlr h, const hmat_kernel* k,
                                    int64_t N, int32_t leaf_size);
/* Same with x / y already on the device (inputs resident in HBM). */
hmat_status hmat_hodlr_build_kernel_device(hmat_hodlr h,
                                           const hmat_kernel* k_device_points,
                                           int64_t N, int32_t leaf_size);
/* Batch of independent systems (no reference counterpart): system s owns
 * rows [off_s, off_s + sizes[s]) of the concatenated points (off_s = sum of
 * the sizes before s) and gets its own bisection tree; the handle then holds
 * one block-diagonal operator whose levels are built, factored and solved
 * for all systems at once.  matvec / solve / solve_transpose take
 * sum(sizes) rows; per-system right-hand-side counts are expressed by zero
 * columns.  Kernel parameters are shared: for per-system Gaussian widths,
 * scale the points of system s by 1/sigma_s and use p0 = 1. */
hmat_status hmat_hodlr_build_kernel_batch_device(hmat_hodlr h,
                                                 const hmat_kernel* k_device_points,
                                                 const int64_t* sizes,
                                                 int32_t nsys, int32_t leaf_size);
/* Build + factor + solve of such a batch in one call: X = A^{-1} B for
 * ncols columns of sum(sizes) rows (device arrays; the handle's refinement
 * target applies, per column).  use_graph = 0: build, factor and solve as
 * the three calls above.  use_graph = 1 (fixed rank, FP64, generated
 * entries): the work takes a fixed schedule without host synchronisation
 * (shifted CholeskyQR3 bases, the refinement count the first call of the
 * shape needed) and is replayed from one HIP graph.  The first call for a
 * shape runs that schedule eagerly and records it; a later call with the
 * same shape (sizes, leaf size, kernel type, point arrays, p0, column
 * count, options) replays it, reading the current point values (rescale
 * them in place for new widths) and the kernel's diagonal shift p1 (for
 * the kernels that have one: Gaussian and the reference kernels).  One
 * download after a replay checks it; a replay whose residual misses the
 * target is refined further eagerly; one whose Cholesky or pivot statuses
 * fail (a basis that needs more CholeskyQR passes for the new values) is
 * redone eagerly and recorded again.  Any other build or load on the handle, or a
 * transposed matvec or solve, drops the graph.  *path (optional): 0 eager,
 * 1 eager and recorded, 2 replayed, 3 replayed and finished eagerly.  No
 * reference counterpart. */
hmat_status hmat_hodlr_batch_solve_device(hmat_hodlr h, const hmat_kernel* k_device_points,
                                          const int64_t* sizes, int32_t nsys, int32_t leaf_size,
                                          const double* B, int64_t ldb, int32_t ncols, double* X,
                                          int64_t ldx, int32_t use_graph, int32_t* path);
/* Why the handle's last hmat_hodlr_batch_solve_device with use_graph = 1
 * neither recorded nor replayed a graph ("" when it did); valid until the
 * next call on the handle. */
const char* hmat_hodlr_batch_note(hmat_hodlr h);
/* Number of doubles of host Omega that build_dense expects. */
int64_t hmat_hodlr_omega_length(const hmat_options* opts, int64_t N,
                                int32_t leaf_size);

/* HODLRMatrixOMP::multiply(const Matrix& x)   hodlr_routines_omp.h:81-135
 *   Y = A_H X for nrhs columns (the reference uses column 0 only). */
hmat_status hmat_hodlr_matvec(hmat_hodlr h, const double* X, int64_t ldx,
                              int32_t nrhs, double* Y, int64_t ldy);
hmat_status hmat_hodlr_matvec_device(hmat_hodlr h, const double* dX,
                                     int64_t ldx, int32_t nrhs, double* dY,
                                     int64_t ldy);

/* Allocation-free hmat_hodlr_matvec_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels as hmat_hodlr_matvec_device (1, 2 or 4 columns
 * per pass; equal results bit for bit up to 4 columns), with their partial
 * products in one fixed device array per GPU of
 * hmat_hodlr_matvec_noalloc_capacity() doubles (32 MiB), held until the
 * process ends, instead of buffers of the handle.  The call allocates
 * nothing and synchronises the handle's stream.  An operator whose partials
 * exceed the array returns HMAT_ERR_ARG.  Calls that overlap on one GPU
 * race for the array (serialise them). */
hmat_status hmat_hodlr_matvec_device_noalloc(hmat_hodlr h, const double* dX, int64_t ldx, int32_t nrhs,
                                             double* dY, int64_t ldy);
int64_t hmat_hodlr_matvec_noalloc_capacity(void);

/* Factor once for HODLRMatrixOMP::solve (hodlr_routines_omp.h:137-240);
 * solve() factors on first use when this was not called. */
hmat_status hmat_hodlr_factor(hmat_hodlr h);
/* HODLRMatrixOMP::solve(const Matrix& b)     hodlr_routines_omp.h:137-145
 *   X = A_H^{-1} B, nrhs columns.  In HMAT_PREC_MIXED the factorisation is
 *   FP32 and the answer is refined in FP64 until ||B - A_H X|| <= rtol ||B||
 *   per column (rtol from hmat_hodlr_set_refinement, default 1e-10; in FP64
 *   refinement is off unless set).  Refinement stops early when a correction
 *   does not reduce the residual; X then holds the iterate with the lowest
 *   residual.  A missed target returns HMAT_ERR_NOT_CONVERGED (X written,
 *   last_refinement tells the residual reached). */
hmat_status hmat_hodlr_solve(hmat_hodlr h, const double* B, int64_t ldb,
                             int32_t nrhs, double* X, int64_t ldx);
hmat_status hmat_hodlr_solve_device(hmat_hodlr h, const double* dB,
                                    int64_t ldb, int32_t nrhs, double* dX,
                                    int64_t ldx);
/* Transposed operator, for adjoints (no reference counterpart; the
 * reference has no transpose solve).  A_H^T has the same HODLR structure
 * with the roles of (U S) and V exchanged; its factorisation is kept beside
 * the one of A_H, so repeated adjoint solves factor once. */
hmat_status hmat_hodlr_solve_transpose(hmat_hodlr h, const double* B, int64_t ldb,
                                       int32_t nrhs, double* X, int64_t ldx);
hmat_status hmat_hodlr_solve_transpose_device(hmat_hodlr h, const double* dB,
                                              int64_t ldb, int32_t nrhs,
                                              double* dX, int64_t ldx);
hmat_status hmat_hodlr_matvec_transpose(hmat_hodlr h, const double* X,
                                        int64_t ldx, int32_t nrhs, double* Y,
                                        int64_t ldy);
hmat_status hmat_hodlr_set_refinement(hmat_hodlr h, double rtol,
                                      int32_t max_iters);
/* Iterations and final max relative residual of the last solve. */
hmat_status hmat_hodlr_last_refinement(hmat_hodlr h, int32_t* iters,
                                       double* resid);

/* Introspection (HODLRMatrixOMP public members N, nodes; HODLRNode fields
 * utils.h:6-29). */
int64_t hmat_hodlr_size(hmat_hodlr h);
int32_t hmat_hodlr_levels(hmat_hodlr h);
/* ranks of the blocks at tree level lvl (1..levels), one per node at that
 * level: rank of the block whose rows are that node. */
hmat_status hmat_hodlr_level_ranks(hmat_hodlr h, int32_t lvl, int32_t* ranks,
                                   int32_t cap, int32_t* count);
/* Dense reconstruction of A_H (checking only; N x N host buffer). */
hmat_status hmat_hodlr_to_dense(hmat_hodlr h, double* A, int64_t lda);
/* Operator files (no reference counterpart: the reference has no
 * serialisation).  save writes the built operator (options, tree sizes,
 * ranks, (U S) and V per level, leaf blocks) to `path`; load replaces the
 * handle's operator and options with the file's.  The factorisation is
 * recomputed on the first solve after a load. */
hmat_status hmat_hodlr_save(hmat_hodlr h, const char* path);
hmat_status hmat_hodlr_load(hmat_hodlr h, const char* path);
/* Solves replay their apply-inverse from a HIP graph recorded once per
 * factorisation and column count (default on); off launches the same
 * kernels one by one.  No reference counterpart. */
hmat_status hmat_hodlr_set_solve_graph(hmat_hodlr h, int32_t on);
/* Apply-inverse as one persistent launch per four columns: every
 * workgroup owns leaves, solves their leaf blocks, then meets the other
 * leaves of each parent through device-memory counters, deepest depth
 * first; it is launched cooperatively (all workgroups resident) and every
 * wait is bounded.  mode 1 (default): fused where it is the faster path (at
 * most four columns and four leaves per resident workgroup), multi-launch
 * otherwise; 0: always the multi-launch apply (leaf GEMM, then three
 * launches per depth, graph-recorded); 2: fused whenever the device holds
 * its grid (tests).  The two agree to rounding (different summation
 * order).  No reference counterpart. */
hmat_status hmat_hodlr_set_fused_solve(hmat_hodlr h, int32_t mode);
/* Diagnostic: spin limit of every wait of the persistent kernels, the
 * HODLR apply-inverse and the HSS matvec (process-wide; 0 restores the
 * default, 2^22).  A wait that exceeds it ends its workgroup and the solve
 * or matvec returns HMAT_ERR_DEVICE. */
void hmat_set_fused_spin_limit(int64_t n);
/* Scratch budget in bytes of one group of tree depths built together (the
 * build stacks a group's sketch bases: three N x k arrays per depth; every
 * launch covers all blocks of the group).  Groups hold at least one depth;
 * 0 restores the default (8e9).  Results agree with any grouping to
 * rounding.  No reference counterpart (the reference recurses per node). */
hmat_status hmat_hodlr_set_group_bytes(hmat_hodlr h, double bytes);
/* Set the HIP stream (hipStream_t) used by the handle; NULL = own stream. */
hmat_status hmat_hodlr_set_stream(hmat_hodlr h, void* stream);

/* ------------------------------------------------------------------- HSS -- */
/* HSSMatrixOMP::HSSMatrixOMP(double tol)      hss_routines_omp.h:19 */
hmat_status hmat_hss_create(hmat_hss* h, const hmat_options* opts);
hmat_status hmat_hss_destroy(hmat_hss h);
/* HSSMatrixOMP::build_from_dense              hss_routines_omp.h:62-149
 *   nested = 0: identity transfers, ranks add up (the reference operator).
 *   nested = 1: compressed transfers (true nested bases); leaf bases from
 *     the N-wide off-diagonal row / column blocks as the reference's
 *     (hss_routines_omp.h:69-95), O(N^2 k / leaf) device memory.
 *   nested = 2: the same nested form from one sample pass per side,
 *     Y = A Omega and Z = A^T Psi with s = rank + oversample columns
 *     (randomized HSS construction, Martinsson 2011): O(N s) device memory,
 *     power_iters does not apply.  Bases, transfers and couplings are new
 *     with respect to the reference (which has no such construction).
 *   omega: optional host Omega, or NULL.  nested = 0/1: blocks in the
 *   serial draw order (reverse node id over leaves, row block then column
 *   block, each N x k; hss_routines.h:57-77), omega_len =
 *   hmat_hss_omega_length.  nested = 2: [Omega | Psi], each N x s
 *   column-major, omega_len = hmat_hss_sampled_omega_length. */
hmat_status hmat_hss_build_dense(hmat_hss h, const double* A, int64_t lda,
                                 int64_t N, int32_t leaf_size, int32_t nested,
                                 const double* omega, int64_t omega_len);
hmat_status hmat_hss_build_kernel(hmat_hss h, const hmat_kernel* k, int64_t N,
                                  int32_t leaf_size, int32_t nested);
int64_t hmat_hss_omega_length(const hmat_options* opts, int64_t N,
                              int32_t leaf_size);
/* Host Omega length of the sampled build (nested = 2): 2 N s doubles,
 * s = min(rank + oversample, N).  No reference counterpart. */
int64_t hmat_hss_sampled_omega_length(const hmat_options* opts, int64_t N);
/* build_kernel with the points (k->x, k->y) already in device memory.
 * No reference counterpart (the reference takes a dense host matrix). */
hmat_status hmat_hss_build_kernel_device(hmat_hss h, const hmat_kernel* k, int64_t N,
                                         int32_t leaf_size, int32_t nested);
/* HSSMatrixOMP::multiply                      hss_routines_omp.h:154-230 */
hmat_status hmat_hss_matvec(hmat_hss h, const double* X, int64_t ldx,
                            int32_t nrhs, double* Y, int64_t ldy);
/* The same on device arrays (inputs resident in HBM; ld >= N). */
hmat_status hmat_hss_matvec_device(hmat_hss h, const double* X, int64_t ldx,
                                   int32_t nrhs, double* Y, int64_t ldy);
hmat_status hmat_hss_factor(hmat_hss h);
/* HSSMatrixOMP::solve                         hss_routines_omp.h:235-358 */
hmat_status hmat_hss_solve(hmat_hss h, const double* B, int64_t ldb,
                           int32_t nrhs, double* X, int64_t ldx);
/* The same on device arrays (inputs resident in HBM; ld >= N). */
hmat_status hmat_hss_solve_device(hmat_hss h, const double* B, int64_t ldb,
                                  int32_t nrhs, double* X, int64_t ldx);
int64_t hmat_hss_size(hmat_hss h);
/* rank_U of the root's children summed (reference: root->rank_U). */
int32_t hmat_hss_root_rank(hmat_hss h);
hmat_status hmat_hss_to_dense(hmat_hss h, double* A, int64_t lda);
/* FP64 iterative refinement of HSS solves through the HSS matvec (no
 * reference counterpart; the reference solves once): with rtol > 0 every
 * solve iterates X += A^{-1}(B - A X) until max_j ||b_j - A x_j|| / ||b_j||
 * <= rtol, max_iters corrections, or a correction that does not reduce the
 * residual (X then goes back to the iterate with the lowest residual).  A
 * missed target returns HMAT_ERR_NOT_CONVERGED with X written.  rtol = 0
 * (the default) turns it off.  last_refinement reports the iterations kept
 * and the final max relative residual of the last solve (-1 when
 * refinement was off). */
hmat_status hmat_hss_set_refinement(hmat_hss h, double rtol, int32_t max_iters);
hmat_status hmat_hss_last_refinement(hmat_hss h, int32_t* iters, double* resid);
/* Solver of hmat_hss_factor / hmat_hss_solve (no reference counterpart for
 * ULV: the reference's solve is the push-through Woodbury recursion,
 * hss_routines_omp.h:247-346, HMAT_HSS_WOODBURY, the default).
 * HMAT_HSS_ULV factors the operator by orthogonal eliminations (Householder
 * transforms and triangular solves; Chandrasekaran, Gu and Pals 2006):
 * backward stable without refinement.  It needs every reduced block to
 * have at most 1024 rows (leaves of at most 1024 points, and the children's
 * ranks summed at every node; the nested builds keep them bounded, the
 * reference-mode build's ranks add up) and returns HMAT_ERR_ARG from the
 * factorisation otherwise.  Changing the solver drops the factorisation;
 * refinement (above) runs on top of either solver. */
typedef enum { HMAT_HSS_WOODBURY = 0, HMAT_HSS_ULV = 1 } hmat_hss_solver;
hmat_status hmat_hss_set_solver(hmat_hss h, int32_t solver);
/* Operator files (no reference counterpart: the reference has no
 * serialisation).  save writes the built operator (options, N, leaf size,
 * nested flag, per-node ranks, transfers, couplings, leaf bases and leaf
 * blocks) to `path`; load replaces the handle's operator and options with
 * the file's after checking that every generator lies inside the stored
 * arrays.  The factorisation is recomputed on the first solve after a
 * load. */
hmat_status hmat_hss_save(hmat_hss h, const char* path);
hmat_status hmat_hss_load(hmat_hss h, const char* path);

#ifdef __cplusplus
}
#endif
#endif /* HMAT_AMD_H */
