// Shared host/device helpers for the gfx950 HODLR / HSS engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/hmat_amd.h"

namespace hmat {

// Error carried from deep inside the engine to the C ABI boundary.
struct Error : std::runtime_error {
  hmat_status code;
  Error(hmat_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess) {
    hmat_status c = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
                        ? HMAT_ERR_OOM : HMAT_ERR_DEVICE;
    throw Error(c, std::string(what) + ": " + hipGetErrorString(e) + " at " + file +
                       ":" + std::to_string(line));
  }
}
#define HIP_CHECK(x) ::hmat::hip_check((x), #x, __FILE__, __LINE__)
#define HMAT_REQUIRE(cond, code, msg) \
  do { if (!(cond)) throw ::hmat::Error((code), (msg)); } while (0)

// Device fills and copies as kernels (no memset or memcpy nodes where a
// caller records the stream into a graph).
void zero_bytes_async(void* p, size_t bytes, hipStream_t s);
void copy_bytes_async(void* dst, const void* src, size_t bytes, hipStream_t s);

// Descriptor record of a build that is replayed from a graph.  While a log
// is attached to a stream's Stager in Persist mode, every staged region
// gets a device region of its own that is never reused (its bytes kept in
// the log), and an upload into a caller's buffer (DevBuf::upload) becomes a
// device-to-device copy from such a region.  In Replay mode the same
// sequence of staging calls returns the same regions and enqueues the same
// copies, without host-to-device traffic, after checking that every array
// holds the bytes recorded; a mismatch marks the log failed.  A stream
// being recorded into a graph thus reads only device memory that stays
// valid and unchanged for every replay.
struct StageLog {
  enum Mode { Persist, Replay };
  Mode mode = Persist;
  std::vector<std::vector<uint8_t>> bytes;  // per region
  std::vector<char*> dev;                   // per region
  std::vector<void*> dst;                   // per region: copy target, or null
  size_t next = 0;                          // replay cursor
  bool failed = false;
  std::vector<char*> chunks;                // device arena
  size_t chunk_off = 0, chunk_cap = 0;
  StageLog() = default;
  StageLog(const StageLog&) = delete;
  StageLog& operator=(const StageLog&) = delete;
  ~StageLog();
  char* alloc(size_t bytes);  // a fresh 256-byte aligned region of the arena
};

// Per-stream staging arena for the small descriptor arrays that every
// level-batched launch reads (segments, tiles, work items): a pinned host
// ring and a device ring of the same size.  put_bytes() copies the host data
// into the pinned ring, enqueues the DMA to the device ring on the stream and
// returns the device address, with no allocation and no host
// synchronisation.  Regions are handed out in order; when the ring is full
// the stream is synchronised once and filling restarts at the front, so no
// pinned byte is rewritten before its DMA has run, and device bytes are only
// rewritten by DMAs ordered after every kernel enqueued before them.  All
// users of one arena must enqueue on its stream, and staged addresses must
// not be recorded into a graph.
class Stager {
 public:
  explicit Stager(hipStream_t s) : s_(s) {}
  ~Stager();
  Stager(const Stager&) = delete;
  Stager& operator=(const Stager&) = delete;
  void* put_bytes(const void* h, size_t bytes);
  // several host arrays in one staged region with one DMA; off[i] receives
  // the byte offset of array i from the returned device address
  char* put_gather(const std::pair<const void*, size_t>* items, int n, size_t* off);
  // DMA `bytes` from `h` (pageable, may be freed on return) to `dst`
  void copy_to(void* dst, const void* h, size_t bytes);
  // attach a descriptor record (null detaches)
  void attach(StageLog* log) { log_ = log; }
  StageLog* log() const { return log_; }

 private:
  size_t reserve(size_t need);  // offset of `need` free bytes
  char* logged(const void* h, size_t bytes, void* dst);
  StageLog* log_ = nullptr;
  hipStream_t s_;
  char* host_ = nullptr;
  char* dev_ = nullptr;
  size_t cap_ = 0, off_ = 0;
  std::vector<std::pair<char*, char*>> retired_;  // outgrown arenas, freed at the next sync
};
Stager& stager(hipStream_t s);
// Drop the stream's arena (after synchronising the stream); called before a
// stream the engine owns is destroyed or replaced.
void stager_release(hipStream_t s);
template <class T>
inline const T* stage(const std::vector<T>& v, hipStream_t s) {
  return v.empty() ? nullptr : static_cast<const T*>(stager(s).put_bytes(v.data(), v.size() * sizeof(T)));
}

// The descriptor arrays of one launch (or one plan), staged together.
class StageGroup {
 public:
  template <class T>
  int add(const std::vector<T>& v) {
    items_.push_back({v.data(), v.size() * sizeof(T)});
    return (int)items_.size() - 1;
  }
  // through the stream's staging arena (valid for work enqueued next)
  void put(hipStream_t s) {
    off_.assign(items_.size(), 0);
    base_ = items_.empty() ? nullptr : stager(s).put_gather(items_.data(), (int)items_.size(), off_.data());
  }
  // into a device buffer owned by the caller (persistent plans), one DMA
  template <class Blob>
  void put_into(Blob& blob, hipStream_t s) {
    off_.assign(items_.size(), 0);
    size_t total = 0;
    for (size_t i = 0; i < items_.size(); ++i) {
      off_[i] = total;
      total += (items_[i].second + 255) & ~(size_t)255;
    }
    std::vector<uint8_t> h(total);
    for (size_t i = 0; i < items_.size(); ++i)
      if (items_[i].second) std::memcpy(h.data() + off_[i], items_[i].first, items_[i].second);
    blob.upload(h.data(), total, s);
    base_ = reinterpret_cast<char*>(blob.get());
  }
  template <class T>
  const T* get(int i) const {
    return items_[i].second ? reinterpret_cast<const T*>(base_ + off_[i]) : nullptr;
  }

 private:
  std::vector<std::pair<const void*, size_t>> items_;
  std::vector<size_t> off_;
  char* base_ = nullptr;
};

// Owning device buffer (typed, 256-byte aligned by hipMalloc).
template <class T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { resize(n); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) HIP_CHECK(hipMalloc(&p_, n * sizeof(T)));
    n_ = n;
  }
  // grow-only resize (keeps the allocation when large enough)
  void reserve(size_t n) { if (n > n_) resize(n); }
  void zero(hipStream_t s) { if (n_) zero_bytes_async(p_, n_ * sizeof(T), s); }
  // asynchronous through the stream's staging arena: `h` may be freed on
  // return, and no host synchronisation happens unless the buffer grows
  void upload(const T* h, size_t n, hipStream_t s) {
    if (n > n_) {
      if (p_) HIP_CHECK(hipStreamSynchronize(s));
      resize(n);
    }
    if (n) stager(s).copy_to(p_, h, n * sizeof(T));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

template <class T>
inline std::vector<T> download(const T* d, size_t n, hipStream_t s) {
  std::vector<T> h(n);
  if (n) HIP_CHECK(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return h;
}

// Opt-in kernel timing: HIP events around selected launches on the stream
// they run on, plus the algorithmic work (flops or bytes) of each launch.
struct ProfEntry {
  int64_t launches = 0;
  double ms = 0.0;
  double work = 0.0;
};
bool prof_enabled();
void prof_set(bool on);
void prof_reset();
void prof_add(const char* name, double ms, double work);
// a launch timed by events a..b on its stream, read when queried
void prof_defer(const char* name, hipEvent_t a, hipEvent_t b, double work);
// per-item counts in a device buffer, summed into sum_name (n items into
// n_name) when queried, after the work enqueued on s so far
void prof_defer_count(const char* sum_name, const char* n_name, DevBuf<int>&& buf, size_t n, hipStream_t s);
bool prof_get(const char* name, ProfEntry* out);

// RAII timer: records start at construction, stop + accumulate at finish().
class ProfScope {
 public:
  ProfScope(const char* name, hipStream_t s, double work) : name_(name), s_(s), work_(work) {
    on_ = prof_enabled();
    if (on_) {
      HIP_CHECK(hipEventCreate(&a_));
      HIP_CHECK(hipEventCreate(&b_));
      HIP_CHECK(hipEventRecord(a_, s_));
    }
  }
  // records the stop event; the registry reads the time when queried, so
  // a profiled launch costs no host wait
  void finish() {
    if (!on_) return;
    HIP_CHECK(hipEventRecord(b_, s_));
    prof_defer(name_, a_, b_, work_);
    on_ = false;
  }
  void cancel() {  // drop this scope's record (untimed work such as autotuning)
    if (!on_) return;
    (void)hipEventDestroy(a_);
    (void)hipEventDestroy(b_);
    on_ = false;
  }
  ~ProfScope() {
    if (on_) {
      (void)hipEventDestroy(a_);
      (void)hipEventDestroy(b_);
    }
  }

 private:
  const char* name_;
  hipStream_t s_;
  double work_;
  bool on_ = false;
  hipEvent_t a_{}, b_{};
};

// Tile autotuner: a process-wide table from a launch's shape class to the
// index of the fastest candidate tiling, filled on first use by timing the
// candidates (off by default; tuned launches may round differently from the
// default tiling).
struct TuneKey {
  int kind, trans, kpad, rows_lg, red_lg, nseg_lg;
  bool operator<(const TuneKey& o) const {
    const int a[6] = {kind, trans, kpad, rows_lg, red_lg, nseg_lg};
    const int b[6] = {o.kind, o.trans, o.kpad, o.rows_lg, o.red_lg, o.nseg_lg};
    for (int i = 0; i < 6; ++i)
      if (a[i] != b[i]) return a[i] < b[i];
    return false;
  }
};
bool autotune_enabled();
void autotune_enable(bool on);
// nested suspension of the autotuner / of profiling timers on this thread
void autotune_suspend(bool on);
void prof_suspend(bool on);
bool tune_lookup(const TuneKey& k, int& choice);
void tune_store(const TuneKey& k, int choice);
size_t tune_entries();

// The option rules of hmat_*_create, also applied to options read from an
// operator file.
inline void check_options(const hmat_options& o) {
  HMAT_REQUIRE(o.rank > 0, HMAT_ERR_ARG, "rank must be positive");
  HMAT_REQUIRE(o.oversample >= 0 && o.power_iters >= 0 && o.exact_below >= 0, HMAT_ERR_ARG,
               "negative oversample / power_iters / exact_below");
  HMAT_REQUIRE(o.precision == HMAT_PREC_FP64 || o.precision == HMAT_PREC_MIXED, HMAT_ERR_ARG,
               "unknown precision");
  HMAT_REQUIRE(o.omega_dist == HMAT_OMEGA_UNIFORM || o.omega_dist == HMAT_OMEGA_GAUSSIAN, HMAT_ERR_ARG,
               "unknown Omega distribution");
  HMAT_REQUIRE(o.fixed_rank == 0 || o.fixed_rank == 1, HMAT_ERR_ARG, "fixed_rank must be 0 or 1");
  HMAT_REQUIRE(std::isfinite(o.tol), HMAT_ERR_ARG, "tolerance is not finite");
}

// %.3e for error messages (std::to_string prints 1e-12 as 0.000000)
inline std::string sci(double v) {
  char b[32];
  std::snprintf(b, sizeof(b), "%.3e", v);
  return b;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int round_up(int a, int b) { return (int)((a + b - 1) / b) * b; }

}  // namespace hmat
