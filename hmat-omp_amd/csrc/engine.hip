// Shared level-batched building blocks: segmented shifted CholeskyQR3.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace hmat {

namespace {
std::mutex g_prof_mu;
bool g_prof_on = false;
std::map<std::string, ProfEntry>& prof_map() {
  static std::map<std::string, ProfEntry> m;
  return m;
}
// Timed launches and counter buffers whose results are read only when the
// statistics are queried, so profiling adds no host wait to the work.
struct ProfPending {
  std::string name;
  hipEvent_t a, b;
  double work;
};
struct CountPending {
  std::string sum_name, n_name;
  hipEvent_t done;
  DevBuf<int> buf;
  size_t n;
};
std::vector<ProfPending>& prof_pending() {
  static std::vector<ProfPending> v;
  return v;
}
std::vector<CountPending>& count_pending() {
  static std::vector<CountPending> v;
  return v;
}
void prof_add_locked(const std::string& name, double ms, double work) {
  auto& e = prof_map()[name];
  e.launches += 1;
  e.ms += ms;
  e.work += work;
}
// Reached from the C entry points hmat_profile_query / hmat_profile_reset,
// so it does not throw: an entry whose events or copy fail is dropped.
void prof_resolve_locked() {
  for (auto& p : prof_pending()) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess)
      prof_add_locked(p.name, ms, p.work);
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  prof_pending().clear();
  for (auto& c : count_pending()) {
    std::vector<int> h(c.n);
    const bool ok = hipEventSynchronize(c.done) == hipSuccess &&
                    (c.n == 0 || hipMemcpy(h.data(), c.buf.get(), c.n * sizeof(int), hipMemcpyDeviceToHost) ==
                                     hipSuccess);
    if (ok) {
      double tot = 0.0;
      for (int v : h) tot += v;
      prof_add_locked(c.sum_name, 0.0, tot);
      prof_add_locked(c.n_name, 0.0, (double)c.n);
    }
    (void)hipEventDestroy(c.done);
  }
  count_pending().clear();
}
}  // namespace

namespace {
std::mutex g_tune_mu;
bool g_tune_on = false;
std::map<TuneKey, int>& tune_map() {
  static std::map<TuneKey, int> m;
  return m;
}
}  // namespace

namespace {
std::mutex g_stage_mu;
std::map<hipStream_t, std::unique_ptr<Stager>>& stage_map() {
  static std::map<hipStream_t, std::unique_ptr<Stager>> m;
  return m;
}
constexpr size_t kStageBytes = (size_t)4 << 20;  // grown on demand
}  // namespace

Stager& stager(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_stage_mu);
  auto& p = stage_map()[s];
  if (!p) p.reset(new Stager(s));
  return *p;
}
void stager_release(hipStream_t s) {
  std::unique_ptr<Stager> p;
  {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    auto it = stage_map().find(s);
    if (it == stage_map().end()) return;
    p = std::move(it->second);
    stage_map().erase(it);
  }
  (void)hipStreamSynchronize(s);
}

Stager::~Stager() {
  (void)hipStreamSynchronize(s_);
  for (auto& r : retired_) (void)hipHostFree(r.first), (void)hipFree(r.second);
  if (host_) (void)hipHostFree(host_);
  if (dev_) (void)hipFree(dev_);
}

size_t Stager::reserve(size_t need) {
  if (need > cap_) {
    // a bigger arena; the old one stays valid for what is already enqueued
    // or staged for the next launch, and is freed at the next synchronisation
    if (host_) retired_.push_back({host_, dev_});
    cap_ = std::max(kStageBytes, 2 * need);
    HIP_CHECK(hipHostMalloc((void**)&host_, cap_, hipHostMallocDefault));
    HIP_CHECK(hipMalloc((void**)&dev_, cap_));
    off_ = 0;
  } else if (off_ + need > cap_) {
    HIP_CHECK(hipStreamSynchronize(s_));
    for (auto& r : retired_) HIP_CHECK(hipHostFree(r.first)), HIP_CHECK(hipFree(r.second));
    retired_.clear();
    off_ = 0;
  }
  const size_t o = off_;
  off_ += need;
  return o;
}

StageLog::~StageLog() {
  for (char* c : chunks) (void)hipFree(c);
}

char* StageLog::alloc(size_t bytes) {
  const size_t need = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
  if (chunks.empty() || chunk_off + need > chunk_cap) {
    chunk_cap = std::max<size_t>((size_t)16 << 20, need);
    char* c = nullptr;
    HIP_CHECK(hipMalloc((void**)&c, chunk_cap));
    chunks.push_back(c);
    chunk_off = 0;
  }
  char* p = chunks.back() + chunk_off;
  chunk_off += need;
  return p;
}

// A staged region (dst null) or an upload into `dst`, through the attached
// log: Persist keeps the bytes and uploads them into a region of the log's
// arena (from the kept copy, so the source may be freed at once), Replay
// checks them against the record; both then copy the region into `dst`.
char* Stager::logged(const void* h, size_t bytes, void* dst) {
  StageLog& L = *log_;
  char* d = nullptr;
  if (L.mode == StageLog::Persist) {
    d = L.alloc(bytes);
    L.bytes.emplace_back(static_cast<const uint8_t*>(h), static_cast<const uint8_t*>(h) + bytes);
    L.dev.push_back(d);
    L.dst.push_back(dst);
    HIP_CHECK(hipMemcpyAsync(d, L.bytes.back().data(), bytes, hipMemcpyHostToDevice, s_));
  } else {
    const size_t i = L.next++;
    // the copy target may have moved since the record (a buffer that grew
    // later in the recorded call): the current one is written
    if (i >= L.dev.size() || L.bytes[i].size() != bytes || (L.dst[i] == nullptr) != (dst == nullptr) ||
        (bytes && std::memcmp(L.bytes[i].data(), h, bytes) != 0)) {
      L.failed = true;
      return i < L.dev.size() ? L.dev[i] : nullptr;
    }
    d = L.dev[i];
  }
  if (dst && bytes) copy_bytes_async(dst, d, bytes, s_);
  return d;
}

void* Stager::put_bytes(const void* h, size_t bytes) {
  if (log_) return logged(h, bytes, nullptr);
  const size_t o = reserve((bytes + 255) & ~(size_t)255);
  std::memcpy(host_ + o, h, bytes);
  HIP_CHECK(hipMemcpyAsync(dev_ + o, host_ + o, bytes, hipMemcpyHostToDevice, s_));
  return dev_ + o;
}

char* Stager::put_gather(const std::pair<const void*, size_t>* items, int n, size_t* off) {
  size_t total = 0;
  for (int i = 0; i < n; ++i) {
    off[i] = total;
    total += (items[i].second + 255) & ~(size_t)255;
  }
  if (total == 0) return nullptr;
  if (log_) {
    std::vector<uint8_t> cat(total, 0);
    for (int i = 0; i < n; ++i)
      if (items[i].second) std::memcpy(cat.data() + off[i], items[i].first, items[i].second);
    return logged(cat.data(), total, nullptr);
  }
  const size_t o = reserve(total);
  for (int i = 0; i < n; ++i)
    if (items[i].second) std::memcpy(host_ + o + off[i], items[i].first, items[i].second);
  HIP_CHECK(hipMemcpyAsync(dev_ + o, host_ + o, total, hipMemcpyHostToDevice, s_));
  return dev_ + o;
}

void Stager::copy_to(void* dst, const void* h, size_t bytes) {
  if (log_) {
    (void)logged(h, bytes, dst);
    return;
  }
  const size_t o = reserve((bytes + 255) & ~(size_t)255);
  std::memcpy(host_ + o, h, bytes);
  HIP_CHECK(hipMemcpyAsync(dst, host_ + o, bytes, hipMemcpyHostToDevice, s_));
}

namespace {
// per thread: a recorded build keeps the default tilings (no timing runs)
// and, while its launches are being recorded, takes no timing events
thread_local int g_tune_suspend = 0, g_prof_suspend = 0;
}  // namespace
void autotune_suspend(bool on) { g_tune_suspend += on ? 1 : -1; }
void prof_suspend(bool on) { g_prof_suspend += on ? 1 : -1; }
bool autotune_enabled() { return g_tune_on && g_tune_suspend == 0; }
void autotune_enable(bool on) { g_tune_on = on; }
bool tune_lookup(const TuneKey& k, int& choice) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = tune_map().find(k);
  if (it == tune_map().end()) return false;
  choice = it->second;
  return true;
}
void tune_store(const TuneKey& k, int choice) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  tune_map()[k] = choice;
}
size_t tune_entries() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return tune_map().size();
}

bool prof_enabled() { return g_prof_on && g_prof_suspend == 0; }
void prof_set(bool on) { g_prof_on = on; }
void prof_reset() {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_resolve_locked();
  prof_map().clear();
}
void prof_add(const char* name, double ms, double work) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_add_locked(name, ms, work);
}
void prof_defer(const char* name, hipEvent_t a, hipEvent_t b, double work) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  // a long profiling session without queries stays bounded
  if (prof_pending().size() >= 4096) prof_resolve_locked();
  prof_pending().push_back(ProfPending{name, a, b, work});
}
void prof_defer_count(const char* sum_name, const char* n_name, DevBuf<int>&& buf, size_t n, hipStream_t s) {
  hipEvent_t e;
  HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIP_CHECK(hipEventRecord(e, s));
  std::lock_guard<std::mutex> lk(g_prof_mu);
  count_pending().push_back(CountPending{sum_name, n_name, e, std::move(buf), n});
}
bool prof_get(const char* name, ProfEntry* out) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_resolve_locked();
  auto it = prof_map().find(name);
  if (it == prof_map().end()) return false;
  *out = it->second;
  return true;
}

bool Orthogonalizer::pass(const std::vector<OrthSeg>& segs, const double* in, double* out,
                          int64_t ld, int kcap, bool shifted, hipStream_t s) {
  const size_t nb = segs.size();
  const int64_t kk = (int64_t)kcap * kcap;
  G_.reserve(nb * kk);
  R_.reserve(nb * kk);
  Rchk_.reserve(nb * kk);
  info_.reserve(nb);
  gerr_.reserve(nb);
  std::vector<SegTN> tn(nb);
  std::vector<SmallMat> sm(nb);
  std::vector<double> coef(nb), shift(nb);
  std::vector<SegNN> nn(nb);
  for (size_t b = 0; b < nb; ++b) {
    const auto& g = segs[b];
    tn[b] = SegTN{g.off, g.off, (int64_t)b * kk, g.len, g.k, g.k, 0};
    sm[b] = SmallMat{(int64_t)b * kk, g.k, kcap};
    // shift 11 (m n + n (n+1)) u ||X||^2 with ||X||_F^2 = trace(G) >= ||X||_2^2
    shift[b] = 11.0 * ((double)g.len * g.k + (double)g.k * (g.k + 1)) * 1.1102230246251565e-16;
    coef[b] = shifted ? shift[b] : 0.0;
    nn[b] = SegNN{g.off, (int64_t)b * kk, g.off, g.len, g.k, g.k, 0};
  }
  kern::seg_gemm_tn(tn, in, ld, in, ld, G_.get(), kcap, 1.0, 0.0, ws_, s);
  // an unshifted pass whose Cholesky breaks down (columns numerically
  // dependent) takes the shifted factor in the same launch
  kern::chol_shifted(sm, G_.get(), R_.get(), coef, info_.get(), gerr_.get(), s,
                     shifted ? std::vector<double>() : shift);
  auto info = download(info_.get(), nb, s);
  auto gerr = download(gerr_.get(), nb, s);
  bool breakdown = false;
  double dev = 0.0;
  for (size_t b = 0; b < nb; ++b) {
    if (info[b] < 0) throw Error(HMAT_ERR_NONFINITE, "orthogonalisation: non-finite Gram matrix");
    if (info[b] > 0) breakdown = true;
    dev = std::max(dev, gerr[b]);
  }
  if (!shifted && dev < kOrthTol) return false;
  if (breakdown) {
    if (shifted) throw Error(HMAT_ERR_NONFINITE, "orthogonalisation: shifted Cholesky broke down");
    return pass(segs, in, out, ld, kcap, true, s);
  }
  kern::seg_trsm_upper(nn, in, ld, R_.get(), kcap, out, ld, s);
  return true;
}

void Orthogonalizer::make_plan(const std::vector<OrthSeg>& segs, int kcap, bool with_acc, hipStream_t s) {
  const size_t nb = segs.size();
  const int64_t kk = (int64_t)kcap * kcap;
  std::vector<SegTN> tn(nb);
  std::vector<SmallMat> sm(nb);
  std::vector<SegNN> nn(nb), acc;
  std::vector<double> coef(nb), zero(nb, 0.0);
  plan_.nmax = 0;
  plan_.panel = 0;
  for (size_t b = 0; b < nb; ++b) {
    const auto& g = segs[b];
    plan_.panel += (double)g.len * g.k;
    tn[b] = SegTN{g.off, g.off, (int64_t)b * kk, g.len, g.k, g.k, 0};
    sm[b] = SmallMat{(int64_t)b * kk, g.k, kcap};
    // shift 11 (m n + n (n+1)) u ||X||^2 with ||X||_F^2 = trace(G) >= ||X||_2^2
    coef[b] = 11.0 * ((double)g.len * g.k + (double)g.k * (g.k + 1)) * 1.1102230246251565e-16;
    nn[b] = SegNN{g.off, (int64_t)b * kk, g.off, g.len, g.k, g.k, 0};
    if (with_acc) acc.push_back(SegNN{(int64_t)b * kk, (int64_t)b * kk, (int64_t)b * kk, g.k, g.k, g.k, 0});
    plan_.nmax = std::max(plan_.nmax, g.k);
  }
  std::vector<kern::TNItem> items;
  std::vector<kern::TNOut> outs;
  plan_.nslots = kern::tn_items(tn, items, outs);
  plan_.nitems = (int)items.size();
  plan_.nouts = (int)outs.size();
  std::vector<kern::NNTile> trt, acct;
  plan_.kmax = kern::trsm_tiles(nn, trt);
  plan_.tr_ntiles = (int)trt.size();
  plan_.acc_mode = with_acc ? kern::nn_tiles(acc, acct) : 0;
  plan_.acc_ntiles = (int)acct.size();
  StageGroup g;
  const int i_tn = g.add(tn), i_it = g.add(items), i_out = g.add(outs), i_sm = g.add(sm), i_on = g.add(coef),
            i_off = g.add(zero), i_tr = g.add(nn), i_trt = g.add(trt), i_acc = g.add(acc), i_acct = g.add(acct);
  g.put_into(plan_.blob, s);
  plan_.tn_segs = g.get<SegTN>(i_tn);
  plan_.tn_items = g.get<kern::TNItem>(i_it);
  plan_.tn_outs = g.get<kern::TNOut>(i_out);
  plan_.mats = g.get<SmallMat>(i_sm);
  plan_.shift_on = g.get<double>(i_on);
  plan_.shift_off = g.get<double>(i_off);
  plan_.tr_segs = g.get<SegNN>(i_tr);
  plan_.tr_tiles = g.get<kern::NNTile>(i_trt);
  plan_.acc_segs = g.get<SegNN>(i_acc);
  plan_.acc_tiles = g.get<kern::NNTile>(i_acct);
  G_.reserve(nb * kk);
  R_.reserve(nb * kk);
  Rchk_.reserve(nb * kk);
  Racc_.reserve(nb * kk);
  if (with_acc) Rtmp_.reserve(nb * kk);
  ws_.part.reserve((size_t)std::max(plan_.nslots, 1) * 4096);
}

// Timed as "orth_gram" / "orth_trsm" when profiling, with their algorithmic
// HBM bytes: the panel read once (Gram), read and written once (TRSM).
void Orthogonalizer::gram(const double* Y, int64_t ld, int kcap, hipStream_t s) {
  ProfScope prof("orth_gram", s, 8.0 * plan_.panel);
  kern::launch_tn_raw<double>(plan_.nitems, plan_.nouts, plan_.tn_segs, plan_.tn_items, plan_.tn_outs, Y, ld, Y,
                              ld, G_.get(), kcap, 1.0, 0.0, ws_.part.get(), s);
  prof.finish();
}

void Orthogonalizer::trsm(const double* X, int64_t ld, const double* R, int kcap, double* Y, hipStream_t s) {
  ProfScope prof("orth_trsm", s, 16.0 * plan_.panel);
  kern::launch_trsm_raw(plan_.kmax, plan_.tr_ntiles, plan_.tr_segs, plan_.tr_tiles, X, ld, R, kcap, Y, ld, s);
  prof.finish();
}

// Racc <- R_ * Racc (upper triangular products)
void Orthogonalizer::acc_product(int kcap, hipStream_t s) {
  kern::launch_nn_raw<double>(plan_.acc_mode, plan_.acc_ntiles, plan_.acc_segs, plan_.acc_tiles, R_.get(), kcap,
                              Racc_.get(), kcap, Rtmp_.get(), kcap, 1.0, 0.0, s);
  std::swap(Racc_, Rtmp_);
}

void Orthogonalizer::run_once(const std::vector<OrthSeg>& segs, double* Y, double* T, int64_t ld, int kcap,
                              hipStream_t s) {
  if (segs.empty()) return;
  const size_t nb = segs.size();
  make_plan(segs, kcap, false, s);
  gerr_.reserve(nb);
  // this pass's Cholesky status goes to its own slice of once_info_, read
  // with the next run()'s status download
  if (once_off_ + nb > once_info_.size()) {
    check_once(s);
    once_info_.reserve(std::max<size_t>(nb, 4096));
  }
  int* info = once_info_.get() + once_off_;
  once_off_ += nb;
  gram(Y, ld, kcap, s);
  kern::launch_chol_shifted(plan_.mats, (int)nb, plan_.nmax, G_.get(), R_.get(), plan_.shift_on, info,
                            gerr_.get(), s);
  if (plan_.kmax <= 64) {
    trsm(Y, ld, R_.get(), kcap, Y, s);  // in place
  } else {
    trsm(Y, ld, R_.get(), kcap, T, s);
    std::vector<SegNN> cp(nb);
    for (size_t b = 0; b < nb; ++b) cp[b] = SegNN{segs[b].off, 0, segs[b].off, segs[b].len, 0, segs[b].k, 0};
    kern::copy_segs(cp, dnn_, T, ld, Y, ld, s);
  }
}

// Statuses of the run_once passes enqueued since the last check: a shifted
// Cholesky that broke down left R unwritten (only non-finite input does
// that, the shift keeps a finite Gram matrix positive definite).
void Orthogonalizer::check_once(hipStream_t s) {
  if (once_off_ == 0) return;
  const auto h = download(once_info_.get(), once_off_, s);
  once_off_ = 0;
  for (int v : h) {
    if (v < 0) throw Error(HMAT_ERR_NONFINITE, "orthogonalisation: non-finite Gram matrix");
    if (v > 0) throw Error(HMAT_ERR_NONFINITE, "orthogonalisation: shifted Cholesky broke down");
  }
}

void Orthogonalizer::run_fixed(const std::vector<OrthSeg>& segs, double* Y, double* T, int64_t ld, int kcap,
                               double* Rout, hipStream_t s) {
  const size_t nb = segs.size();
  const int64_t kk = (int64_t)kcap * kcap;
  make_plan(segs, kcap, Rout != nullptr, s);
  const size_t c = fixed_call_++;
  if (c == fixed_info_.size()) {
    fixed_info_.emplace_back();
    fixed_gerr_.emplace_back();
    fixed_nb_.push_back(0);
    fixed_passes_.push_back(0);
    fixed_dev_.push_back(0.0);
  }
  HMAT_REQUIRE(fixed_choose_ || fixed_nb_[c] == nb, HMAT_ERR_STATE, "recorded orthogonalisation differs");
  fixed_nb_[c] = nb;
  DevBuf<int>& st = fixed_info_[c];
  DevBuf<double>& ge = fixed_gerr_[c];
  st.reserve(2 * nb);
  ge.reserve(nb);
  const int nm = (int)nb;
  // bases of at most 64 columns are solved in place; wider ones alternate
  // between Y and T and are copied back at the end
  const bool inplace = plan_.kmax <= 64;
  // shifted pass: R_1 in Racc_
  gram(Y, ld, kcap, s);
  kern::launch_chol_shifted(plan_.mats, nm, plan_.nmax, G_.get(), Racc_.get(), plan_.shift_on, st.get(), ge.get(), s);
  double* cur = inplace ? Y : T;
  trsm(Y, ld, Racc_.get(), kcap, cur, s);
  for (int p = 0;; ++p) {
    // the Gram matrix of the current basis: its deviation from I checks it,
    // its (fallback-shifted where needed) factor is the next pass's
    gram(cur, ld, kcap, s);
    kern::launch_chol_shifted(plan_.mats, nm, plan_.nmax, G_.get(), R_.get(), plan_.shift_off, st.get() + nb,
                              ge.get(), s, plan_.shift_on);
    bool more;
    if (fixed_choose_) {
      const auto hi = download(st.get(), 2 * nb, s);
      const auto hg = download(ge.get(), nb, s);
      double dev = 0.0;
      bool clean = true;
      for (size_t b = 0; b < nb; ++b) {
        if (hi[b] < 0 || hi[nb + b] < 0) throw Error(HMAT_ERR_NONFINITE, "orthogonalisation: non-finite Gram matrix");
        if (hi[b] > 0) throw Error(HMAT_ERR_NONFINITE, "orthogonalisation: shifted Cholesky broke down");
        if (hi[nb + b] != 0) clean = false;
        dev = std::max(dev, hg[b]);
      }
      more = !(clean && dev < kOrthTol) && p + 2 < kOrthMaxPasses;
      if (!more) {
        fixed_passes_[c] = p;
        fixed_dev_[c] = dev;
      }
    } else {
      more = p < fixed_passes_[c];
    }
    if (!more) break;
    double* nxt = inplace ? cur : (cur == Y ? T : Y);
    trsm(cur, ld, R_.get(), kcap, nxt, s);
    if (Rout) acc_product(kcap, s);
    cur = nxt;
  }
  if (!inplace) {
    // always enqueued (a copy onto itself when the basis is in Y already), so
    // that the staged descriptors do not depend on the number of passes
    std::vector<SegNN> cp(nb);
    for (size_t b = 0; b < nb; ++b) cp[b] = SegNN{segs[b].off, 0, segs[b].off, segs[b].len, 0, segs[b].k, 0};
    kern::copy_segs(cp, dnn_, cur, ld, Y, ld, s);
  }
  if (Rout) copy_bytes_async(Rout, Racc_.get(), nb * kk * sizeof(double), s);
}

std::string Orthogonalizer::check_fixed(size_t ncalls, hipStream_t s) {
  if (ncalls > fixed_info_.size()) return "orthogonalisation: statuses of a recorded build are missing";
  for (size_t c = 0; c < ncalls; ++c) {
    const size_t nb = fixed_nb_[c];
    const auto h = download(fixed_info_[c].get(), 2 * nb, s);
    const auto g = download(fixed_gerr_[c].get(), nb, s);
    const double bound = std::max(kOrthTol, 10.0 * fixed_dev_[c]);
    for (size_t b = 0; b < nb; ++b) {
      if (h[b] < 0 || h[nb + b] < 0 || !std::isfinite(g[b])) return "orthogonalisation: non-finite Gram matrix";
      if (h[b] > 0) return "orthogonalisation: shifted Cholesky broke down";
      if (!(g[b] <= bound))
        return "orthogonalisation: basis deviates from orthonormal by " + sci(g[b]) + " (recorded " +
               sci(fixed_dev_[c]) + ")";
    }
  }
  return std::string();
}

void Orthogonalizer::run(const std::vector<OrthSeg>& segs, double* Y, double* T, int64_t ld,
                         int kcap, double* Rout, hipStream_t s) {
  if (segs.empty()) return;
  if (fixed_) {
    run_fixed(segs, Y, T, ld, kcap, Rout, s);
    return;
  }
  const size_t nb = segs.size();
  const int64_t kk = (int64_t)kcap * kcap;
  make_plan(segs, kcap, Rout != nullptr, s);
  const size_t info_bytes = (3 * nb * sizeof(int) + 255) & ~(size_t)255;
  chk_.reserve(info_bytes + 3 * nb * sizeof(double));
  int* info = reinterpret_cast<int*>(chk_.get());
  double* gerr = reinterpret_cast<double*>(chk_.get() + info_bytes);
  const int nm = (int)nb;
  // Enqueued back to back, with one download of all three status arrays:
  // shifted pass Y -> T (R_1 in Racc_), unshifted pass T -> Y (R_2 in R_;
  // shifted instead for an entry whose unshifted Cholesky breaks down),
  // and the Gram matrix of the result with its deviation from I and its
  // Cholesky factor (Rchk_, the third pass's when one is needed).
  gram(Y, ld, kcap, s);
  kern::launch_chol_shifted(plan_.mats, nm, plan_.nmax, G_.get(), Racc_.get(), plan_.shift_on, info, gerr, s);
  trsm(Y, ld, Racc_.get(), kcap, T, s);
  gram(T, ld, kcap, s);
  kern::launch_chol_shifted(plan_.mats, nm, plan_.nmax, G_.get(), R_.get(), plan_.shift_off, info + nb, gerr + nb, s,
                            plan_.shift_on);
  trsm(T, ld, R_.get(), kcap, Y, s);
  gram(Y, ld, kcap, s);
  kern::launch_chol_shifted(plan_.mats, nm, plan_.nmax, G_.get(), Rchk_.get(), plan_.shift_off, info + 2 * nb,
                            gerr + 2 * nb, s, plan_.shift_on);
  const auto h = download(chk_.get(), info_bytes + 3 * nb * sizeof(double), s);
  check_once(s);  // the stream is drained: this download costs no extra wait
  const int* hi = reinterpret_cast<const int*>(h.data());
  const double* hg = reinterpret_cast<const double*>(h.data() + info_bytes);
  bool breakdown2 = false, again = false;
  for (size_t b = 0; b < nb; ++b) {
    if (hi[b] < 0) throw Error(HMAT_ERR_NONFINITE, "orthogonalisation: non-finite Gram matrix");
    if (hi[b] > 0) throw Error(HMAT_ERR_NONFINITE, "orthogonalisation: shifted Cholesky broke down");
    if (hi[nb + b] != 0) breakdown2 = true;
    if (hi[2 * nb + b] != 0 || !(hg[2 * nb + b] < kOrthTol)) again = true;
  }
  double* src = Y;
  double* dst = T;
  if (breakdown2) {
    // the unshifted pass broke down: continue from T = Y_in R_1^{-1}
    // with a shifted pass
    src = T;
    dst = Y;
    again = true;
  } else if (Rout) {
    acc_product(kcap, s);
  }
  int p = 2;
  if (again && !breakdown2) {
    bool factored = true;
    for (size_t b = 0; b < nb; ++b)
      if (hi[2 * nb + b] != 0) factored = false;
    if (factored) {
      // the check's Cholesky factor of Y (shifted only where the unshifted
      // one broke down) is the next pass's factor: apply it without
      // forming the same Gram matrix again
      std::swap(R_, Rchk_);
      trsm(src, ld, R_.get(), kcap, dst, s);
      if (Rout) acc_product(kcap, s);
      std::swap(src, dst);
      p = 3;
    }
  }
  // rare: more passes, each checked, until the Gram matrix is I to kOrthTol
  bool orthonormal = !again;
  for (; again && p < kOrthMaxPasses; ++p) {
    if (!pass(segs, src, dst, ld, kcap, breakdown2 && src == T && p == 2, s)) {
      orthonormal = true;
      break;
    }
    if (Rout) acc_product(kcap, s);
    std::swap(src, dst);
  }
  // CholeskyQR stalls on blocks whose columns span many orders of magnitude
  // below the shift (Gaussian blocks whose singular values fall to
  // underflow, an exactly zero row): the columns it cannot resolve have norms
  // ~1e-12 of the block's and carry nothing measurable into Q R, which stays
  // an exact factorisation.  The last pass's basis is kept and the event is
  // counted ("orth_unconverged", calls with a segment left unresolved).
  if (!orthonormal) prof_add("orth_unconverged", 0.0, 1.0);
  if (src != Y) {
    std::vector<SegNN> cp(nb);
    for (size_t b = 0; b < nb; ++b) cp[b] = SegNN{segs[b].off, 0, segs[b].off, segs[b].len, 0, segs[b].k, 0};
    kern::copy_segs(cp, dnn_, src, ld, Y, ld, s);
  }
  if (Rout) HIP_CHECK(hipMemcpyAsync(Rout, Racc_.get(), nb * kk * sizeof(double), hipMemcpyDeviceToDevice, s));
}

}  // namespace hmat
