// extern "C" boundary of the engine (declared in include/hmat_amd.h).
#include <cstring>
#include <string>

#include "hodlr.hpp"
#include "hss.hpp"

using namespace hmat;

struct hmat_hodlr_s {
  HodlrEngine eng;
  DevBuf<double> a, x, y, px, py;
  explicit hmat_hodlr_s(const hmat_options& o) : eng(o) {}
};
struct hmat_hss_s {
  HssEngine eng;
  DevBuf<double> a, x, y;
  explicit hmat_hss_s(const hmat_options& o) : eng(o) {}
};

namespace {
thread_local std::string g_err;

template <class F>
hmat_status guard(F&& f) {
  try {
    f();
    g_err.clear();
    return HMAT_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return HMAT_ERR_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return HMAT_ERR_DEVICE;
  }
}

void check_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
    throw Error(HMAT_ERR_DEVICE, "no HIP device available");
}

void check_opts(const hmat_options* o) {
  HMAT_REQUIRE(o != nullptr, HMAT_ERR_ARG, "options pointer is NULL");
  check_options(*o);
}

// Pack a host column-major (ld) block into a dense device N x nc array.
void upload_cols(DevBuf<double>& d, const double* h, int64_t ld, int64_t n, int64_t nc, hipStream_t s) {
  d.reserve((size_t)(n * nc));
  HIP_CHECK(hipMemcpy2DAsync(d.get(), n * sizeof(double), h, ld * sizeof(double), n * sizeof(double),
                             nc, hipMemcpyHostToDevice, s));
}
void download_cols(double* h, int64_t ld, const DevBuf<double>& d, int64_t n, int64_t nc, hipStream_t s) {
  HIP_CHECK(hipMemcpy2DAsync(h, ld * sizeof(double), d.get(), n * sizeof(double), n * sizeof(double),
                             nc, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}
}  // namespace

extern "C" {

const char* hmat_last_error(void) { return g_err.c_str(); }
const char* hmat_version(void) { return "hmat-amd 0.1 (gfx950)"; }
int32_t hmat_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

void hmat_profile_enable(int32_t on) { prof_set(on != 0); }
void hmat_profile_reset(void) { prof_reset(); }
void hmat_autotune_enable(int32_t on) { autotune_enable(on != 0); }
int64_t hmat_autotune_entries(void) { return (int64_t)tune_entries(); }
int32_t hmat_profile_query(const char* name, int64_t* launches, double* ms, double* work) {
  ProfEntry e;
  if (!name || !prof_get(name, &e)) return 0;
  if (launches) *launches = e.launches;
  if (ms) *ms = e.ms;
  if (work) *work = e.work;
  return 1;
}

void hmat_options_reference(hmat_options* o, double tol) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->tol = tol;
  o->rank = 64;
  o->fixed_rank = 0;
  o->oversample = 10;
  o->power_iters = 0;
  o->exact_below = 256;
  o->precision = HMAT_PREC_FP64;
  o->omega_dist = HMAT_OMEGA_UNIFORM;
  o->seed = 0x5EEDull;
}

// ------------------------------------------------------------------ HODLR --
hmat_status hmat_hodlr_create(hmat_hodlr* h, const hmat_options* opts) {
  return guard([&] {
    HMAT_REQUIRE(h != nullptr, HMAT_ERR_ARG, "handle pointer is NULL");
    check_opts(opts);
    check_device();
    *h = new hmat_hodlr_s(*opts);
  });
}

hmat_status hmat_hodlr_destroy(hmat_hodlr h) {
  return guard([&] { delete h; });
}

int64_t hmat_hodlr_omega_length(const hmat_options* opts, int64_t N, int32_t leaf) {
  if (!opts || N <= 0 || leaf <= 0) return -1;
  int64_t t = 0;
  for (auto& e : HodlrEngine::omega_schedule(*opts, N, leaf)) t += e.first * e.second;
  return t;
}

static hmat_status hodlr_build_common(hmat_hodlr h, const double* dA, int64_t lda, int64_t N,
                                      int32_t leaf, const double* omega, int64_t omega_len) {
  GenDesc g;
  g.kind = GenKind::Dense;
  g.A = dA;
  g.lda = lda;
  h->eng.build(g, N, leaf, omega, omega_len);
  return HMAT_OK;
}

hmat_status hmat_hodlr_build_dense(hmat_hodlr h, const double* A, int64_t lda, int64_t N,
                                   int32_t leaf, const double* omega, int64_t omega_len) {
  return guard([&] {
    HMAT_REQUIRE(h && A && N > 0 && lda >= N && leaf > 0, HMAT_ERR_ARG, "bad build_dense arguments");
    upload_cols(h->a, A, lda, N, N, h->eng.stream);
    hodlr_build_common(h, h->a.get(), N, N, leaf, omega, omega_len);
    h->a.release();
  });
}

hmat_status hmat_hodlr_build_dense_device(hmat_hodlr h, const double* dA, int64_t lda, int64_t N,
                                          int32_t leaf, const double* omega, int64_t omega_len) {
  return guard([&] {
    HMAT_REQUIRE(h && dA && N > 0 && lda >= N && leaf > 0, HMAT_ERR_ARG, "bad build_dense arguments");
    hodlr_build_common(h, dA, lda, N, leaf, omega, omega_len);
  });
}

hmat_status hmat_hodlr_build_kernel(hmat_hodlr h, const hmat_kernel* k, int64_t N, int32_t leaf) {
  return guard([&] {
    HMAT_REQUIRE(h && k && k->x && N > 0 && leaf > 0, HMAT_ERR_ARG, "bad build_kernel arguments");
    HMAT_REQUIRE(k->type != HMAT_KERNEL_CAUCHY || k->y, HMAT_ERR_ARG, "Cauchy kernel needs y points");
    upload_cols(h->px, k->x, N, N, 1, h->eng.stream);
    if (k->y) upload_cols(h->py, k->y, N, N, 1, h->eng.stream);
    GenDesc g = gen_from_kernel(*k, h->px.get(), k->y ? h->py.get() : nullptr);
    h->eng.build(g, N, leaf, nullptr, 0);
  });
}

hmat_status hmat_hodlr_build_kernel_device(hmat_hodlr h, const hmat_kernel* k, int64_t N,
                                           int32_t leaf) {
  return guard([&] {
    HMAT_REQUIRE(h && k && k->x && N > 0 && leaf > 0, HMAT_ERR_ARG, "bad build_kernel arguments");
    HMAT_REQUIRE(k->type != HMAT_KERNEL_CAUCHY || k->y, HMAT_ERR_ARG, "Cauchy kernel needs y points");
    GenDesc g = gen_from_kernel(*k, k->x, k->y);
    h->eng.build(g, N, leaf, nullptr, 0);
  });
}

hmat_status hmat_kernel_matvec_device(const hmat_kernel* k, int64_t N, const double* dX, int64_t ldx,
                                      int32_t nrhs, double* dY, int64_t ldy) {
  return guard([&] {
    HMAT_REQUIRE(k && k->x && dX && dY && N > 0 && nrhs > 0 && ldx >= N && ldy >= N, HMAT_ERR_ARG,
                 "bad kernel_matvec arguments");
    check_device();
    GenDesc g = gen_from_kernel(*k, k->x, k->y);
    hipStream_t s;
    HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    kern::Workspace ws;
    DevBuf<SegSk> ds;
    for (int c0 = 0; c0 < nrhs; c0 += 64) {
      const int nc = std::min(64, nrhs - c0);
      std::vector<SegSk> seg{SegSk{0, 0, 0, (int64_t)c0 * ldx, (int64_t)c0 * ldy, (int32_t)N, (int32_t)N, 0, nc}};
      kern::sketch(g, false, seg, dX, ldx, dY, ldy, nc, ws, s);
    }
    HIP_CHECK(hipStreamSynchronize(s));
    stager_release(s);
    HIP_CHECK(hipStreamDestroy(s));
  });
}

hmat_status hmat_hodlr_matvec(hmat_hodlr h, const double* X, int64_t ldx, int32_t nrhs, double* Y,
                              int64_t ldy) {
  return guard([&] {
    HMAT_REQUIRE(h && X && Y && nrhs > 0, HMAT_ERR_ARG, "bad matvec arguments");
    const int64_t N = h->eng.N;
    HMAT_REQUIRE(h->eng.built, HMAT_ERR_STATE, "matvec before build");
    HMAT_REQUIRE(ldx >= N && ldy >= N, HMAT_ERR_ARG, "leading dimension below N");
    upload_cols(h->x, X, ldx, N, nrhs, h->eng.stream);
    h->y.reserve((size_t)(N * nrhs));
    h->eng.matvec(h->x.get(), N, nrhs, h->y.get(), N);
    download_cols(Y, ldy, h->y, N, nrhs, h->eng.stream);
  });
}

hmat_status hmat_hodlr_matvec_device(hmat_hodlr h, const double* dX, int64_t ldx, int32_t nrhs,
                                     double* dY, int64_t ldy) {
  return guard([&] {
    HMAT_REQUIRE(h && dX && dY && nrhs > 0, HMAT_ERR_ARG, "bad matvec arguments");
    h->eng.matvec(dX, ldx, nrhs, dY, ldy);
  });
}

hmat_status hmat_hodlr_matvec_device_noalloc(hmat_hodlr h, const double* dX, int64_t ldx, int32_t nrhs,
                                             double* dY, int64_t ldy) {
  return guard([&] {
    HMAT_REQUIRE(h && dX && dY && nrhs > 0, HMAT_ERR_ARG, "bad matvec arguments");
    hodlr_matvec_noalloc(h->eng, dX, ldx, nrhs, dY, ldy);
  });
}

int64_t hmat_hodlr_matvec_noalloc_capacity(void) { return (int64_t)hodlr_matvec_noalloc_capacity(); }

hmat_status hmat_hodlr_factor(hmat_hodlr h) {
  return guard([&] {
    HMAT_REQUIRE(h, HMAT_ERR_ARG, "NULL handle");
    h->eng.factor();
  });
}

hmat_status hmat_hodlr_solve(hmat_hodlr h, const double* B, int64_t ldb, int32_t nrhs, double* X,
                             int64_t ldx) {
  return guard([&] {
    HMAT_REQUIRE(h && B && X && nrhs > 0, HMAT_ERR_ARG, "bad solve arguments");
    const int64_t N = h->eng.N;
    HMAT_REQUIRE(h->eng.built, HMAT_ERR_STATE, "solve before build");
    HMAT_REQUIRE(ldb >= N && ldx >= N, HMAT_ERR_ARG, "leading dimension below N");
    upload_cols(h->x, B, ldb, N, nrhs, h->eng.stream);
    h->y.reserve((size_t)(N * nrhs));
    h->eng.solve(h->x.get(), N, nrhs, h->y.get(), N);
    download_cols(X, ldx, h->y, N, nrhs, h->eng.stream);
  });
}

hmat_status hmat_hodlr_solve_device(hmat_hodlr h, const double* dB, int64_t ldb, int32_t nrhs,
                                    double* dX, int64_t ldx) {
  return guard([&] {
    HMAT_REQUIRE(h && dB && dX && nrhs > 0, HMAT_ERR_ARG, "bad solve arguments");
    h->eng.solve(dB, ldb, nrhs, dX, ldx);
  });
}

hmat_status hmat_hodlr_set_refinement(hmat_hodlr h, double rtol, int32_t max_iters) {
  return guard([&] {
    HMAT_REQUIRE(h && rtol > 0 && max_iters >= 0, HMAT_ERR_ARG, "bad refinement arguments");
    h->eng.refine_rtol = rtol;
    h->eng.refine_max = max_iters;
  });
}

hmat_status hmat_hodlr_last_refinement(hmat_hodlr h, int32_t* iters, double* resid) {
  return guard([&] {
    HMAT_REQUIRE(h, HMAT_ERR_ARG, "NULL handle");
    if (iters) *iters = h->eng.last_iters;
    if (resid) *resid = h->eng.last_resid;
  });
}

hmat_status hmat_hodlr_solve_transpose(hmat_hodlr h, const double* B, int64_t ldb, int32_t nrhs, double* X,
                                       int64_t ldx) {
  return guard([&] {
    HMAT_REQUIRE(h && B && X && nrhs > 0, HMAT_ERR_ARG, "bad solve_transpose arguments");
    const int64_t N = h->eng.N;
    HMAT_REQUIRE(h->eng.built, HMAT_ERR_STATE, "solve before build");
    HMAT_REQUIRE(ldb >= N && ldx >= N, HMAT_ERR_ARG, "leading dimension below N");
    upload_cols(h->x, B, ldb, N, nrhs, h->eng.stream);
    h->y.reserve((size_t)(N * nrhs));
    h->eng.solve_transpose(h->x.get(), N, nrhs, h->y.get(), N);
    download_cols(X, ldx, h->y, N, nrhs, h->eng.stream);
  });
}

hmat_status hmat_hodlr_solve_transpose_device(hmat_hodlr h, const double* dB, int64_t ldb, int32_t nrhs,
                                              double* dX, int64_t ldx) {
  return guard([&] {
    HMAT_REQUIRE(h && dB && dX && nrhs > 0, HMAT_ERR_ARG, "bad solve_transpose arguments");
    HMAT_REQUIRE(ldb >= h->eng.N && ldx >= h->eng.N, HMAT_ERR_ARG, "leading dimension below N");
    h->eng.solve_transpose(dB, ldb, nrhs, dX, ldx);
  });
}

hmat_status hmat_hodlr_matvec_transpose(hmat_hodlr h, const double* X, int64_t ldx, int32_t nrhs, double* Y,
                                        int64_t ldy) {
  return guard([&] {
    HMAT_REQUIRE(h && X && Y && nrhs > 0, HMAT_ERR_ARG, "bad matvec_transpose arguments");
    const int64_t N = h->eng.N;
    HMAT_REQUIRE(h->eng.built, HMAT_ERR_STATE, "matvec before build");
    HMAT_REQUIRE(ldx >= N && ldy >= N, HMAT_ERR_ARG, "leading dimension below N");
    upload_cols(h->x, X, ldx, N, nrhs, h->eng.stream);
    h->y.reserve((size_t)(N * nrhs));
    h->eng.matvec_transpose(h->x.get(), N, nrhs, h->y.get(), N);
    download_cols(Y, ldy, h->y, N, nrhs, h->eng.stream);
  });
}

hmat_status hmat_hodlr_build_kernel_batch_device(hmat_hodlr h, const hmat_kernel* k, const int64_t* sizes,
                                                 int32_t nsys, int32_t leaf) {
  return guard([&] {
    HMAT_REQUIRE(h && k && k->x && sizes && nsys > 0 && leaf > 0, HMAT_ERR_ARG,
                 "bad build_kernel_batch arguments");
    HMAT_REQUIRE(k->type != HMAT_KERNEL_CAUCHY || k->y, HMAT_ERR_ARG, "Cauchy kernel needs y points");
    HMAT_REQUIRE(k->type != HMAT_KERNEL_GAUSSIAN_DSIGMA, HMAT_ERR_ARG, "derivative kernel cannot be compressed");
    std::vector<int64_t> sz(sizes, sizes + nsys);
    GenDesc g = gen_from_kernel(*k, k->x, k->y);
    h->eng.build_forest(g, sz, leaf);
  });
}

hmat_status hmat_hodlr_batch_solve_device(hmat_hodlr h, const hmat_kernel* k, const int64_t* sizes, int32_t nsys,
                                          int32_t leaf, const double* B, int64_t ldb, int32_t ncols, double* X,
                                          int64_t ldx, int32_t use_graph, int32_t* path) {
  return guard([&] {
    HMAT_REQUIRE(h && k && k->x && sizes && nsys > 0 && leaf > 0 && B && X && ncols > 0, HMAT_ERR_ARG,
                 "bad batch_solve arguments");
    HMAT_REQUIRE(k->type != HMAT_KERNEL_CAUCHY || k->y, HMAT_ERR_ARG, "Cauchy kernel needs y points");
    HMAT_REQUIRE(k->type != HMAT_KERNEL_GAUSSIAN_DSIGMA, HMAT_ERR_ARG, "derivative kernel cannot be compressed");
    std::vector<int64_t> sz(sizes, sizes + nsys);
    for (int64_t n : sz) HMAT_REQUIRE(n > 0, HMAT_ERR_ARG, "every system size must be positive");
    GenDesc g = gen_from_kernel(*k, k->x, k->y);
    const int p = h->eng.batch_solve(g, sz, leaf, B, ldb, ncols, X, ldx, use_graph != 0);
    if (path) *path = p;
  });
}

const char* hmat_hodlr_batch_note(hmat_hodlr h) { return h ? h->eng.batch_note.c_str() : ""; }

hmat_status hmat_hodlr_save(hmat_hodlr h, const char* path) {
  return guard([&] {
    HMAT_REQUIRE(h && path, HMAT_ERR_ARG, "bad save arguments");
    h->eng.save(path);
  });
}

hmat_status hmat_hodlr_load(hmat_hodlr h, const char* path) {
  return guard([&] {
    HMAT_REQUIRE(h && path, HMAT_ERR_ARG, "bad load arguments");
    h->eng.load(path);
  });
}

hmat_status hmat_hodlr_set_solve_graph(hmat_hodlr h, int32_t on) {
  return guard([&] {
    HMAT_REQUIRE(h, HMAT_ERR_ARG, "NULL handle");
    h->eng.solve_graph = on != 0;
  });
}

hmat_status hmat_hodlr_set_fused_solve(hmat_hodlr h, int32_t mode) {
  return guard([&] {
    HMAT_REQUIRE(h, HMAT_ERR_ARG, "NULL handle");
    HMAT_REQUIRE(mode >= 0 && mode <= 2, HMAT_ERR_ARG, "fused solve mode must be 0, 1 or 2");
    h->eng.fused_solve = mode;
  });
}

void hmat_set_fused_spin_limit(int64_t n) {
  kern::apply_fused_set_spin_limit(n);
  kern::hss_mv_set_spin_limit(n);
}

hmat_status hmat_hodlr_set_group_bytes(hmat_hodlr h, double bytes) {
  return guard([&] {
    HMAT_REQUIRE(h && bytes >= 0.0, HMAT_ERR_ARG, "bad group scratch budget");
    h->eng.group_bytes = bytes > 0.0 ? bytes : kGroupScratchBytes;
  });
}

int64_t hmat_hodlr_size(hmat_hodlr h) { return h ? h->eng.N : -1; }
int32_t hmat_hodlr_levels(hmat_hodlr h) { return h && h->eng.built ? h->eng.tree.max_depth : -1; }

hmat_status hmat_hodlr_level_ranks(hmat_hodlr h, int32_t lvl, int32_t* ranks, int32_t cap,
                                   int32_t* count) {
  return guard([&] {
    HMAT_REQUIRE(h && h->eng.built, HMAT_ERR_STATE, "no built operator");
    HMAT_REQUIRE(lvl >= 1 && lvl <= h->eng.tree.max_depth, HMAT_ERR_ARG, "level out of range");
    const auto& r = h->eng.lv[lvl].rank;
    if (count) *count = (int32_t)r.size();
    if (ranks)
      for (int32_t i = 0; i < cap && i < (int32_t)r.size(); ++i) ranks[i] = r[i];
  });
}

hmat_status hmat_hodlr_to_dense(hmat_hodlr h, double* A, int64_t lda) {
  return guard([&] {
    HMAT_REQUIRE(h && A && lda >= h->eng.N, HMAT_ERR_ARG, "bad to_dense arguments");
    h->eng.to_dense_host(A, lda);
  });
}

hmat_status hmat_hodlr_set_stream(hmat_hodlr h, void* stream) {
  return guard([&] {
    HMAT_REQUIRE(h, HMAT_ERR_ARG, "NULL handle");
    if (h->eng.stream) stager_release(h->eng.stream);
    if (h->eng.own_stream && h->eng.stream) HIP_CHECK(hipStreamDestroy(h->eng.stream));
    if (stream) {
      h->eng.stream = (hipStream_t)stream;
      h->eng.own_stream = false;
    } else {
      HIP_CHECK(hipStreamCreateWithFlags(&h->eng.stream, hipStreamNonBlocking));
      h->eng.own_stream = true;
    }
  });
}

// -------------------------------------------------------------------- HSS --
hmat_status hmat_hss_create(hmat_hss* h, const hmat_options* opts) {
  return guard([&] {
    HMAT_REQUIRE(h != nullptr, HMAT_ERR_ARG, "handle pointer is NULL");
    check_opts(opts);
    check_device();
    *h = new hmat_hss_s(*opts);
  });
}

hmat_status hmat_hss_destroy(hmat_hss h) {
  return guard([&] { delete h; });
}

int64_t hmat_hss_omega_length(const hmat_options* opts, int64_t N, int32_t leaf) {
  if (!opts || N <= 0 || leaf <= 0) return -1;
  return HssEngine::omega_length(*opts, N, leaf);
}

hmat_status hmat_hss_build_dense(hmat_hss h, const double* A, int64_t lda, int64_t N, int32_t leaf,
                                 int32_t nested, const double* omega, int64_t omega_len) {
  return guard([&] {
    HMAT_REQUIRE(h && A && N > 0 && lda >= N && leaf > 0, HMAT_ERR_ARG, "bad build_dense arguments");
    upload_cols(h->a, A, lda, N, N, h->eng.stream);
    GenDesc g;
    g.kind = GenKind::Dense;
    g.A = h->a.get();
    g.lda = N;
    h->eng.build(g, N, leaf, nested, omega, omega_len);
    h->a.release();
  });
}

hmat_status hmat_hss_build_kernel(hmat_hss h, const hmat_kernel* k, int64_t N, int32_t leaf,
                                  int32_t nested) {
  return guard([&] {
    HMAT_REQUIRE(h && k && k->x && N > 0 && leaf > 0, HMAT_ERR_ARG, "bad build_kernel arguments");
    HMAT_REQUIRE(k->type != HMAT_KERNEL_CAUCHY || k->y, HMAT_ERR_ARG, "Cauchy kernel needs y points");
    upload_cols(h->eng.px, k->x, N, N, 1, h->eng.stream);
    if (k->y) upload_cols(h->eng.py, k->y, N, N, 1, h->eng.stream);
    GenDesc g = gen_from_kernel(*k, h->eng.px.get(), k->y ? h->eng.py.get() : nullptr);
    h->eng.build(g, N, leaf, nested, nullptr, 0);
  });
}

hmat_status hmat_hss_build_kernel_device(hmat_hss h, const hmat_kernel* k, int64_t N, int32_t leaf,
                                         int32_t nested) {
  return guard([&] {
    HMAT_REQUIRE(h && k && k->x && N > 0 && leaf > 0, HMAT_ERR_ARG, "bad build_kernel arguments");
    HMAT_REQUIRE(k->type != HMAT_KERNEL_CAUCHY || k->y, HMAT_ERR_ARG, "Cauchy kernel needs y points");
    GenDesc g = gen_from_kernel(*k, k->x, k->y);
    h->eng.build(g, N, leaf, nested, nullptr, 0);
  });
}

int64_t hmat_hss_sampled_omega_length(const hmat_options* opts, int64_t N) {
  if (!opts || N <= 0) return -1;
  return HssEngine::sampled_omega_length(*opts, N);
}

hmat_status hmat_hss_matvec(hmat_hss h, const double* X, int64_t ldx, int32_t nrhs, double* Y,
                            int64_t ldy) {
  return guard([&] {
    HMAT_REQUIRE(h && X && Y && nrhs > 0, HMAT_ERR_ARG, "bad matvec arguments");
    const int64_t N = h->eng.N;
    HMAT_REQUIRE(h->eng.built, HMAT_ERR_STATE, "matvec before build");
    HMAT_REQUIRE(ldx >= N && ldy >= N, HMAT_ERR_ARG, "leading dimension below N");
    upload_cols(h->x, X, ldx, N, nrhs, h->eng.stream);
    h->y.reserve((size_t)(N * nrhs));
    h->eng.matvec(h->x.get(), N, nrhs, h->y.get(), N);
    download_cols(Y, ldy, h->y, N, nrhs, h->eng.stream);
  });
}

hmat_status hmat_hss_matvec_device(hmat_hss h, const double* X, int64_t ldx, int32_t nrhs, double* Y,
                                   int64_t ldy) {
  return guard([&] {
    HMAT_REQUIRE(h && X && Y && nrhs > 0, HMAT_ERR_ARG, "bad matvec arguments");
    h->eng.matvec(X, ldx, nrhs, Y, ldy);
  });
}

hmat_status hmat_hss_solve_device(hmat_hss h, const double* B, int64_t ldb, int32_t nrhs, double* X,
                                  int64_t ldx) {
  return guard([&] {
    HMAT_REQUIRE(h && B && X && nrhs > 0, HMAT_ERR_ARG, "bad solve arguments");
    h->eng.solve(B, ldb, nrhs, X, ldx);
    HIP_CHECK(hipStreamSynchronize(h->eng.stream));
  });
}

hmat_status hmat_hss_factor(hmat_hss h) {
  return guard([&] {
    HMAT_REQUIRE(h, HMAT_ERR_ARG, "NULL handle");
    h->eng.factor();
  });
}

hmat_status hmat_hss_solve(hmat_hss h, const double* B, int64_t ldb, int32_t nrhs, double* X,
                           int64_t ldx) {
  return guard([&] {
    HMAT_REQUIRE(h && B && X && nrhs > 0, HMAT_ERR_ARG, "bad solve arguments");
    const int64_t N = h->eng.N;
    HMAT_REQUIRE(h->eng.built, HMAT_ERR_STATE, "solve before build");
    HMAT_REQUIRE(ldb >= N && ldx >= N, HMAT_ERR_ARG, "leading dimension below N");
    upload_cols(h->x, B, ldb, N, nrhs, h->eng.stream);
    h->y.reserve((size_t)(N * nrhs));
    h->eng.solve(h->x.get(), N, nrhs, h->y.get(), N);
    download_cols(X, ldx, h->y, N, nrhs, h->eng.stream);
  });
}

int64_t hmat_hss_size(hmat_hss h) { return h ? h->eng.N : -1; }
int32_t hmat_hss_root_rank(hmat_hss h) { return h && h->eng.built ? h->eng.root_rank() : -1; }

hmat_status hmat_hss_set_refinement(hmat_hss h, double rtol, int32_t max_iters) {
  return guard([&] {
    HMAT_REQUIRE(h && rtol >= 0 && max_iters >= 0, HMAT_ERR_ARG, "bad refinement arguments");
    h->eng.refine_rtol = rtol;
    h->eng.refine_max = max_iters;
  });
}

hmat_status hmat_hss_set_solver(hmat_hss h, int32_t solver) {
  return guard([&] {
    HMAT_REQUIRE(h, HMAT_ERR_ARG, "NULL handle");
    HMAT_REQUIRE(solver == HMAT_HSS_WOODBURY || solver == HMAT_HSS_ULV, HMAT_ERR_ARG, "unknown HSS solver");
    if (h->eng.solver != solver) h->eng.factored = false;
    h->eng.solver = solver;
  });
}

hmat_status hmat_hss_last_refinement(hmat_hss h, int32_t* iters, double* resid) {
  return guard([&] {
    HMAT_REQUIRE(h, HMAT_ERR_ARG, "NULL handle");
    if (iters) *iters = h->eng.last_iters;
    if (resid) *resid = h->eng.last_resid;
  });
}

hmat_status hmat_hss_save(hmat_hss h, const char* path) {
  return guard([&] {
    HMAT_REQUIRE(h && path, HMAT_ERR_ARG, "bad save arguments");
    h->eng.save(path);
  });
}

hmat_status hmat_hss_load(hmat_hss h, const char* path) {
  return guard([&] {
    HMAT_REQUIRE(h && path, HMAT_ERR_ARG, "bad load arguments");
    h->eng.load(path);
  });
}

hmat_status hmat_hss_to_dense(hmat_hss h, double* A, int64_t lda) {
  return guard([&] {
    HMAT_REQUIRE(h && A && lda >= h->eng.N, HMAT_ERR_ARG, "bad to_dense arguments");
    h->eng.to_dense_host(A, lda);
  });
}

}  // extern "C"
