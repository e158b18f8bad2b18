// Shared level-batched building blocks: segmented shifted CholeskyQR3.
#include <algorithm>
#include <cmath>

#include "engine.hpp"

#include <map>
#include <mutex>
#include <string>

namespace hmat {

namespace {
std::mutex g_prof_mu;
bool g_prof_on = false;
std::map<std::string, ProfEntry>& prof_map() {
  static std::map<std::string, ProfEntry> m;
  return m;
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

bool autotune_enabled() { return g_tune_on; }
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

bool prof_enabled() { return g_prof_on; }
void prof_set(bool on) { g_prof_on = on; }
void prof_reset() {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  prof_map().clear();
}
void prof_add(const char* name, double ms, double work) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  auto& e = prof_map()[name];
  e.launches += 1;
  e.ms += ms;
  e.work += work;
}
bool prof_get(const char* name, ProfEntry* out) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
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
  Ri_.reserve(nb * kk);
  info_.reserve(nb);
  gerr_.reserve(nb);
  std::vector<SegTN> tn(nb);
  std::vector<SmallMat> sm(nb);
  std::vector<double> coef(nb);
  std::vector<SegNN> nn(nb);
  for (size_t b = 0; b < nb; ++b) {
    const auto& g = segs[b];
    tn[b] = SegTN{g.off, g.off, (int64_t)b * kk, g.len, g.k, g.k, 0};
    sm[b] = SmallMat{(int64_t)b * kk, g.k, kcap};
    // shift 11 (m n + n (n+1)) u ||X||^2 with ||X||_F^2 = trace(G) >= ||X||_2^2
    coef[b] = shifted ? 11.0 * ((double)g.len * g.k + (double)g.k * (g.k + 1)) * 1.1102230246251565e-16
                      : 0.0;
    nn[b] = SegNN{g.off, (int64_t)b * kk, g.off, g.len, g.k, g.k, 0};
  }
  kern::seg_gemm_tn(tn, dtn_, in, ld, in, ld, G_.get(), kcap, 1.0, 0.0, ws_, s);
  kern::chol_rinv(sm, dsm_, G_.get(), R_.get(), Ri_.get(), coef, shift_, info_.get(), gerr_.get(), s);
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
  kern::seg_trsm_upper(nn, dnn_, in, ld, R_.get(), kcap, out, ld, s);
  return true;
}

void Orthogonalizer::run(const std::vector<OrthSeg>& segs, double* Y, double* T, int64_t ld,
                         int kcap, double* Rout, hipStream_t s) {
  if (segs.empty()) return;
  const size_t nb = segs.size();
  const int64_t kk = (int64_t)kcap * kcap;
  std::vector<SegNN> acc(nb);
  for (size_t b = 0; b < nb; ++b)
    acc[b] = SegNN{(int64_t)b * kk, (int64_t)b * kk, (int64_t)b * kk, segs[b].k, segs[b].k, segs[b].k, 0};
  if (Rout) {
    Racc_.reserve(nb * kk);
    Rtmp_.reserve(nb * kk);
  }
  double* src = Y;
  double* dst = T;
  for (int p = 0; p < kOrthMaxPasses; ++p) {
    if (!pass(segs, src, dst, ld, kcap, p == 0, s)) break;
    if (Rout) {
      if (p == 0) {
        HIP_CHECK(hipMemcpyAsync(Racc_.get(), R_.get(), nb * kk * sizeof(double), hipMemcpyDeviceToDevice, s));
      } else {
        // Racc <- R_p * Racc (upper triangular products)
        kern::seg_gemm_nn(acc, dnn_, R_.get(), kcap, Racc_.get(), kcap, Rtmp_.get(), kcap, 1.0, 0.0, s);
        std::swap(Racc_, Rtmp_);
      }
    }
    std::swap(src, dst);
  }
  if (src != Y) {
    std::vector<SegNN> cp(nb);
    for (size_t b = 0; b < nb; ++b) cp[b] = SegNN{segs[b].off, 0, segs[b].off, segs[b].len, 0, segs[b].k, 0};
    kern::copy_segs(cp, dnn_, src, ld, Y, ld, s);
  }
  if (Rout) HIP_CHECK(hipMemcpyAsync(Rout, Racc_.get(), nb * kk * sizeof(double), hipMemcpyDeviceToDevice, s));
}

}  // namespace hmat
