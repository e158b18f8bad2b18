// One-launch HSS matvec, y = A_HSS x for one right-hand side.
//
// Reference: HSSMatrixOMP::multiply (include/hss_routines_omp.h:154-230):
// up-sweep g = V^T x at the leaves and g_p = W_p^T [g_a; g_b] above,
// down-sweep f_a = B_ab g_b + R_a f_p, f_b = B_ba g_a + R_b f_p, and
// y = D x + U f at the leaves.  The multi-launch form takes one launch per
// depth and sweep direction (19 at N = 2^18, leaf 256) plus three leaf
// passes, each launch a few microseconds of latency-bound work.  Here one
// persistent cooperative launch does all of it:
//   1. every workgroup, for its leaves: y = D x and g = V^T x in one pass
//      over the leaf's rows (x staged in LDS), g published with a flag;
//   2. the workgroups take the sweep's node tasks from a queue in
//      dependency order (up-sweep deepest first, then down-sweep from the
//      root), each task loading its transfer or coupling matrices into
//      registers before it waits for the flags of its inputs;
//   3. every workgroup, for its leaves: y += U f once f is flagged.
// A task only waits for tasks earlier in the queue or for phase 1, and
// every workgroup is resident (cooperative launch), so the queue drains.
// Every wait is bounded: a wait past its spin limit raises `err` and the
// workgroup returns, and so do the workgroups that then wait on it.
// Vectors and flags handed between workgroups live in uncached memory and
// are written and read with relaxed agent-scope atomics; a writer's stores
// complete (s_waitcnt vmcnt(0)) before the barrier that precedes its flag.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>

#include "hss_mv.hpp"

namespace hmat {
namespace kern {

namespace {
constexpr int kMvWG = 256;
constexpr int kMvWaves = kMvWG / 64;
constexpr int kMvMaxRpt = 4;  // leaves of up to 1024 rows
constexpr int kMvMaxRank = 64;
constexpr int kMvPre = 32;  // entries of R per thread loaded before a down task's last wait
std::atomic<int64_t> g_mv_spin{(int64_t)1 << 22};

using gp = const __attribute__((address_space(1))) double*;

__device__ __forceinline__ void st_rel(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_rel(const double* p) {
  return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void stores_done() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Thread 0 waits until the flags f0, f1 and f2 (those not null) are set,
// the workgroup at the barrier.  False when the spin limit was exceeded or
// err was raised.
__device__ bool mv_wait(int* f0, int* f1, int* f2, int* err, int64_t limit, int* sh) {
  if (threadIdx.x == 0) {
    int ok = 1;
    for (int q = 0; q < 3 && ok; ++q) {
      int* f = q == 0 ? f0 : (q == 1 ? f1 : f2);
      if (f == nullptr) continue;
      int64_t it = 0;
      while (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (++it > limit || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    *sh = ok;
  }
  __syncthreads();
  const bool ok = *sh != 0;
  __syncthreads();
  return ok;
}

__device__ __forceinline__ void mv_flag(int* f) {
  stores_done();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(f, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int S>
__device__ __forceinline__ void mv_halving(double (&v)[16], int lane) {
  if constexpr (S >= 1) {
    const bool upper = (lane & S) != 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
      const double send = upper ? v[i] : v[i + S];
      const double keep = upper ? v[i + S] : v[i];
      v[i] = keep + __shfl_xor(send, S);
    }
    mv_halving<S / 2>(v, lane);
  }
}
// 16 values per lane in; lane l ends with the wave total of v[l & 15]
__device__ __forceinline__ double mv_reduce16(double (&v)[16]) {
  const int lane = threadIdx.x & 63;
  mv_halving<8>(v, lane);
  double s = v[0];
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s;
}

// Up task: g_p = W^T [g_a; g_b].  Wave w takes outputs w*16 .. w*16+15,
// lane l the reduction indices l and l + 64.
__device__ bool mv_up(const HmvArgs& a, const HmvTask& T, int* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nk = T.ra + T.rb;
  gp W = (gp)(a.pool + T.o1);
  double wv[2][16];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = lane + 64 * h;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int c = w * 16 + u;
      wv[h][u] = (k < nk && c < T.rp) ? W[k + (int64_t)c * T.ld1] : 0.0;
    }
  }
  if (!mv_wait(a.gflag + T.a, a.gflag + T.b, nullptr, a.err, a.spin_limit, sh)) return false;
  double gk[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = lane + 64 * h;
    gk[h] = k < T.ra ? ld_rel(a.G + (int64_t)T.a * a.ldv + k)
                     : (k < nk ? ld_rel(a.G + (int64_t)T.b * a.ldv + (k - T.ra)) : 0.0);
  }
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) v[u] = fma(wv[1][u], gk[1], wv[0][u] * gk[0]);
  const double s = mv_reduce16(v);
  const int c = w * 16 + (lane & 15);
  if (lane < 16 && c < T.rp) st_rel(a.G + (int64_t)T.p * a.ldv + c, s);
  mv_flag(a.gflag + T.p);
  return true;
}

// Down task: rows i < m = ra + rb of [f_a; f_b], thread (i, s) for the
// k-slice s of S = min(4, 256 / m).  The coupling part (B_ab g_b or B_ba
// g_a) is formed as soon as the children's g are flagged, which is long
// before f_p; then the thread's slice of R is loaded into registers, and
// only f_p itself is read after the last wait.  vs: 3 x 64 doubles, red:
// 256 doubles of LDS.
__device__ bool mv_down(const HmvArgs& a, const HmvTask& T, double* vs, double* red, int* sh) {
  const int tid = threadIdx.x;
  const int m = T.ra + T.rb;
  const int S = min(4, kMvWG / max(m, 1));
  const int i = tid % max(m, 1), s = tid / max(m, 1);
  const bool live = m > 0 && s < S;
  if (!mv_wait(a.gflag + T.a, a.gflag + T.b, nullptr, a.err, a.spin_limit, sh)) return false;
  if (tid < T.gb) vs[tid] = ld_rel(a.G + (int64_t)T.b * a.ldv + tid);
  if (tid >= 64 && tid - 64 < T.ga) vs[tid] = ld_rel(a.G + (int64_t)T.a * a.ldv + (tid - 64));
  __syncthreads();
  double acc = 0.0;
  if (live) {
    const bool top = i < T.ra;
    gp M = top ? (gp)(a.pool + T.o1) + i : (gp)(a.pool + T.o2) + (i - T.ra);
    const int nk = top ? T.gb : T.ga, voff = top ? 0 : 64;
    const int kq = (nk + S - 1) / S, k0 = s * kq, k1 = min(nk, k0 + kq);
    for (int k = k0; k < k1; k += 16) {
      const int n = min(16, k1 - k);
      double e[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) e[u] = M[(int64_t)(k + min(u, n - 1)) * T.ld1];  // in bounds
#pragma unroll
      for (int u = 0; u < 16; ++u)
        if (u < n) acc = fma(e[u], vs[voff + k + u], acc);
    }
  }
  if (T.rp > 0) {
    // R's slice (at most 32 entries: rp <= 64 and S >= 2) before f_p's wait
    int k0 = 0, nr = 0;
    gp R = nullptr;
    if (live) {
      const int kq = (T.rp + S - 1) / S;
      k0 = s * kq;
      nr = max(0, min(T.rp, k0 + kq) - k0);
      R = (gp)(a.pool + T.o3) + i;
    }
    double pf[kMvPre];
#pragma unroll
    for (int u = 0; u < kMvPre; ++u) pf[u] = u < nr ? R[(int64_t)(k0 + u) * T.ld2] : 0.0;
    if (!mv_wait(a.fflag + T.p, nullptr, nullptr, a.err, a.spin_limit, sh)) return false;
    if (tid >= 128 && tid - 128 < T.rp) vs[tid] = ld_rel(a.F + (int64_t)T.p * a.ldv + (tid - 128));
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kMvPre; ++u)
      if (u < nr) acc = fma(pf[u], vs[128 + k0 + u], acc);
  }
  red[tid] = live ? acc : 0.0;
  __syncthreads();
  if (tid < m) {
    double f = 0.0;
    for (int u = 0; u < S; ++u) f += red[u * m + tid];
    if (tid < T.ra)
      st_rel(a.F + (int64_t)T.a * a.ldv + tid, f);
    else
      st_rel(a.F + (int64_t)T.b * a.ldv + (tid - T.ra), f);
  }
  stores_done();
  __syncthreads();
  if (tid == 0) {
    __hip_atomic_store(a.fflag + T.a, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.fflag + T.b, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

__global__ __launch_bounds__(kMvWG, 4) void k_hss_matvec(HmvArgs a) {
  __shared__ double xs[kMvMaxRpt * kMvWG];
  __shared__ double red[kMvWG];
  __shared__ double vs[3 * kMvMaxRank];
  __shared__ int sh;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int rpt = (int)((a.lm + kMvWG - 1) / kMvWG);

  // 1. y = D x, g = V^T x over the owned leaves
  for (int li = blockIdx.x; li < a.nleaf; li += gridDim.x) {
    const HmvLeaf lf = a.leaves[li];
    for (int e = tid; e < lf.len; e += kMvWG) xs[e] = a.X[lf.start + e];
    __syncthreads();
    gp D = (gp)(a.D + (int64_t)li * a.lm * a.lm);
    for (int t = 0; t < rpt; ++t) {
      const int r = tid + t * kMvWG;
      if (r >= lf.len) break;
      double acc = 0.0;
      int k = 0;
      for (; k + 16 <= lf.len; k += 16) {
        double d[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) d[u] = D[r + (int64_t)(k + u) * a.lm];
#pragma unroll
        for (int u = 0; u < 16; ++u) acc = fma(d[u], xs[k + u], acc);
      }
      for (; k < lf.len; ++k) acc = fma(D[r + (int64_t)k * a.lm], xs[k], acc);
      a.Y[lf.start + r] = acc;
    }
    gp V = (gp)a.V + lf.start;
    for (int kb = 0; kb < lf.rV; kb += 16) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = 0.0;
      for (int t = 0; t < rpt; ++t) {
        const int r = tid + t * kMvWG;
        const bool live = r < lf.len;
        const double xr = live ? xs[r] : 0.0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          const int c = kb + u;
          const double vc = (live && c < lf.rV) ? V[(live ? r : 0) + (int64_t)c * a.N] : 0.0;
          v[u] = fma(vc, xr, v[u]);
        }
      }
      const double sum = mv_reduce16(v);
      if (lane < 16) red[w * kMvMaxRank + kb + lane] = sum;
    }
    __syncthreads();
    for (int c = tid; c < lf.rV; c += kMvWG) {
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < kMvWaves; ++u) s += red[u * kMvMaxRank + c];
      st_rel(a.G + (int64_t)lf.id * a.ldv + c, s);
    }
    mv_flag(a.gflag + lf.id);
    __syncthreads();  // xs and red are reused by the next leaf
  }

  // 2. the sweeps' node tasks, in queue order
  for (;;) {
    if (tid == 0) sh = __hip_atomic_fetch_add(a.counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int t = sh;
    __syncthreads();
    if (t >= a.ntasks) break;
    const HmvTask T = a.tasks[t];
    const bool ok = T.kind == 0 ? mv_up(a, T, &sh) : mv_down(a, T, vs, red, &sh);
    if (!ok) return;
    __syncthreads();
  }

  // 3. y += U f over the owned leaves (the first 16 columns of U of the
  // first row block loaded before the wait)
  for (int li = blockIdx.x; li < a.nleaf; li += gridDim.x) {
    const HmvLeaf lf = a.leaves[li];
    gp U = (gp)a.U + lf.start;
    const int np = (tid < lf.len) ? min(16, lf.rU) : 0;
    double up[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) up[u] = u < np ? U[tid + (int64_t)u * a.N] : 0.0;
    if (!mv_wait(a.fflag + lf.id, nullptr, nullptr, a.err, a.spin_limit, &sh)) return;
    if (tid < lf.rU) vs[tid] = ld_rel(a.F + (int64_t)lf.id * a.ldv + tid);
    __syncthreads();
    for (int t = 0; t < rpt; ++t) {
      const int r = tid + t * kMvWG;
      if (r >= lf.len) break;
      double acc = 0.0;
      int k = 0;
      if (t == 0) {
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (u < np) acc = fma(up[u], vs[u], acc);
        k = np;
      }
      for (; k + 8 <= lf.rU; k += 8) {
        double e[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) e[u] = U[r + (int64_t)(k + u) * a.N];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = fma(e[u], vs[k + u], acc);
      }
      for (; k < lf.rU; ++k) acc = fma(U[r + (int64_t)k * a.N], vs[k], acc);
      a.Y[lf.start + r] += acc;
    }
    __syncthreads();
  }
}

__global__ void k_mv_zero(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}
}  // namespace

void hss_mv_set_spin_limit(int64_t n) { g_mv_spin = n > 0 ? n : ((int64_t)1 << 22); }

HmvPlan::~HmvPlan() {
  if (err_host) (void)hipHostFree(err_host);
}

int hss_mv_grid(const HmvPlan& p) {
  if (!p.ok || p.lm > (int64_t)kMvWG * kMvMaxRpt || p.ldv > kMvMaxRank) return 0;
  int dev = 0, ncu = 0, coop = 0, nb = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) return 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_hss_matvec, kMvWG, 0));
  return std::min(p.nleaf, nb * ncu);
}

bool hss_mv_launch(HmvPlan& p, const double* D, const double* U, const double* V, const double* pool, int64_t N,
                   const double* X, double* Y, hipStream_t s) {
  const int grid = p.grid;
  if (!p.ok || grid <= 0) return false;
  if (!p.err_host) {
    HIP_CHECK(hipHostMalloc((void**)&p.err_host, sizeof(int), hipHostMallocDefault));
    *p.err_host = 0;
  }
  HmvArgs a;
  a.leaves = p.leaves.get();
  a.tasks = p.tasks.get();
  a.D = D;
  a.U = U;
  a.V = V;
  a.pool = pool;
  a.X = X;
  a.Y = Y;
  a.G = p.vec.as<double>();
  a.F = a.G + (int64_t)p.nn * p.ldv;
  int* sy = p.sync.as<int>();
  a.gflag = sy;
  a.fflag = sy + p.nn;
  a.counter = sy + 2 * p.nn;
  a.err = sy + 2 * p.nn + 1;
  a.N = N;
  a.lm = p.lm;
  a.spin_limit = g_mv_spin.load();
  a.nleaf = p.nleaf;
  a.ntasks = p.ntasks;
  a.ldv = p.ldv;
  a.pad_ = 0;
  // flags, counter and err start at zero (a kernel, not a memset)
  const int nz = 2 * p.nn + 2;
  k_mv_zero<<<(unsigned)std::max(1, std::min(64, (nz + 255) / 256)), 256, 0, s>>>(sy, nz);
  HIP_CHECK(hipGetLastError());
  void* args[] = {&a};
  const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_hss_matvec), dim3(grid),
                                                  dim3(kMvWG), args, 0, s);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

void hss_mv_finish(HmvPlan& p, hipStream_t s) {
  int* sy = p.sync.as<int>();
  HIP_CHECK(hipMemcpyAsync(p.err_host, sy + 2 * p.nn + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (*p.err_host) throw Error(HMAT_ERR_DEVICE, "HSS matvec: a wait of the persistent kernel exceeded its spin limit");
}

}  // namespace kern
}  // namespace hmat
