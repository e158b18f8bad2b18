// One-launch HSS matvec, Y = A_HSS X for up to 16 right-hand sides per
// launch (column groups of 16, 8, 4, 2, 1).
//
// Reference: HSSMatrixOMP::multiply (include/hss_routines_omp.h:154-230):
// up-sweep g = V^T x at the leaves and g_p = W_p^T [g_a; g_b] above,
// down-sweep f_a = B_ab g_b + R_a f_p, f_b = B_ba g_a + R_b f_p, and
// y = D x + U f at the leaves.  The multi-launch form takes one launch per
// depth and sweep direction (19 at N = 2^18, leaf 256) plus three leaf
// passes, each launch a few microseconds of latency-bound work.  Here one
// persistent cooperative launch does all of it:
//   1. every workgroup, for its leaves: Y = D X and g = V^T X in one pass
//      over the leaf's rows (X staged in LDS), g published with a flag;
//   2. the workgroups take the sweep's node tasks from a queue in
//      dependency order (up-sweep deepest first, then down-sweep from the
//      root), each task loading its transfer or coupling matrices into
//      registers before it waits for the flags of its inputs;
//   3. every workgroup, for its leaves: Y += U F once F is flagged.
// A task's matrices stay in registers for all the launch's columns.
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

// Per-node vectors of a launch over NR columns: node id's g (or f) is an
// (ldv x NR) block at G + id * ldv * NR, entry (c, q) at c * NR + q.

// Up task: g_p = W^T [g_a; g_b].  Wave w takes outputs w*16 .. w*16+15,
// lane l the reduction indices l and l + 64; W stays in registers for the
// NR columns, whose [g_a; g_b] are staged in LDS (vs: 128 x NR).
template <int NR>
__device__ bool mv_up(const HmvArgs& a, const HmvTask& T, double* vs, int* sh) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nk = T.ra + T.rb;
  const int64_t vst = (int64_t)a.ldv * NR;
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
  for (int e = tid; e < nk * NR; e += kMvWG) {
    const int k = e / NR, q = e % NR;
    vs[e] = k < T.ra ? ld_rel(a.G + T.a * vst + k * NR + q) : ld_rel(a.G + T.b * vst + (k - T.ra) * NR + q);
  }
  __syncthreads();
  const int c = w * 16 + (lane & 15);
  for (int q = 0; q < NR; ++q) {
    const double g0 = lane < nk ? vs[lane * NR + q] : 0.0;
    const double g1 = lane + 64 < nk ? vs[(lane + 64) * NR + q] : 0.0;
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = fma(wv[1][u], g1, wv[0][u] * g0);
    const double s = mv_reduce16(v);
    if (lane < 16 && c < T.rp) st_rel(a.G + T.p * vst + c * NR + q, s);
  }
  mv_flag(a.gflag + T.p);
  return true;
}

// Down task: rows i < m = ra + rb of [f_a; f_b], thread (i, s) for the
// k-slice s of S = min(4, 256 / m).  The coupling part (B_ab g_b or B_ba
// g_a) is formed as soon as the children's g are flagged, which is long
// before f_p; then the thread's slice of R is loaded into registers, and
// only f_p itself is read after the last wait.  vs: 3 x 64 x NR doubles
// (g_b, g_a, f_p), red: 256 x NR doubles of LDS.
template <int NR>
__device__ bool mv_down(const HmvArgs& a, const HmvTask& T, double* vs, double* red, int* sh) {
  const int tid = threadIdx.x;
  const int m = T.ra + T.rb;
  const int S = min(4, kMvWG / max(m, 1));
  const int i = tid % max(m, 1), s = tid / max(m, 1);
  const bool live = m > 0 && s < S;
  const int64_t vst = (int64_t)a.ldv * NR;
  if (!mv_wait(a.gflag + T.a, a.gflag + T.b, nullptr, a.err, a.spin_limit, sh)) return false;
  for (int e = tid; e < 64 * NR; e += kMvWG) {
    const int k = e / NR, q = e % NR;
    if (k < T.gb) vs[e] = ld_rel(a.G + T.b * vst + k * NR + q);
    if (k < T.ga) vs[64 * NR + e] = ld_rel(a.G + T.a * vst + k * NR + q);
  }
  __syncthreads();
  double acc[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) acc[q] = 0.0;
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
        if (u < n)
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = fma(e[u], vs[(voff + k + u) * NR + q], acc[q]);
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
    for (int e = tid; e < T.rp * NR; e += kMvWG) vs[128 * NR + e] = ld_rel(a.F + T.p * vst + e);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kMvPre; ++u)
      if (u < nr)
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = fma(pf[u], vs[(128 + k0 + u) * NR + q], acc[q]);
  }
  if (live)
#pragma unroll
    for (int q = 0; q < NR; ++q) red[(s * m + i) * NR + q] = acc[q];
  __syncthreads();
  for (int e = tid; e < m * NR; e += kMvWG) {
    const int r = e / NR, q = e % NR;
    double f = 0.0;
    for (int u = 0; u < S; ++u) f += red[(u * m + r) * NR + q];
    if (r < T.ra)
      st_rel(a.F + T.a * vst + r * NR + q, f);
    else
      st_rel(a.F + T.b * vst + (r - T.ra) * NR + q, f);
  }
  stores_done();
  __syncthreads();
  if (tid == 0) {
    __hip_atomic_store(a.fflag + T.a, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.fflag + T.b, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

// Dynamic LDS of the launch over NR columns (doubles): phase 1 stages the
// leaf's X (lm x NR) and the waves' partial sums of V^T X (4 x 64 x NR);
// the tasks use 3 x 64 x NR vector entries and 256 x NR partial sums.
inline int64_t mv_lds_doubles(int64_t lm, int nr) { return (int64_t)nr * std::max<int64_t>(lm + 256, 448); }

template <int NR>
__global__ __launch_bounds__(kMvWG, NR == 1 ? 4 : 2) void k_hss_matvec(HmvArgs a) {
  extern __shared__ __align__(16) double sm[];
  __shared__ int sh;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int rpt = (int)((a.lm + kMvWG - 1) / kMvWG);
  const int64_t vst = (int64_t)a.ldv * NR;
  double* xs = sm;                       // phase 1: lm x NR
  double* red = sm + a.lm * NR;          // phase 1: 4 x 64 x NR
  double* vs = sm;                       // tasks, phase 3: 3 x 64 x NR
  double* red2 = sm + 3 * 64 * NR;       // tasks: 256 x NR
  constexpr int KG = 16 / NR;            // columns of V per step of V^T X

  // 1. Y = D X, g = V^T X over the owned leaves
  for (int li = blockIdx.x; li < a.nleaf; li += gridDim.x) {
    const HmvLeaf lf = a.leaves[li];
    for (int e = tid; e < lf.len * NR; e += kMvWG) {
      const int r = e / NR, q = e % NR;
      xs[e] = a.X[lf.start + r + (int64_t)q * a.ldx];
    }
    __syncthreads();
    gp D = (gp)(a.D + (int64_t)li * a.lm * a.lm);
    for (int t = 0; t < rpt; ++t) {
      const int r = tid + t * kMvWG;
      if (r >= lf.len) break;
      double acc[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = 0.0;
      int k = 0;
      for (; k + 16 <= lf.len; k += 16) {
        double d[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) d[u] = D[r + (int64_t)(k + u) * a.lm];
#pragma unroll
        for (int u = 0; u < 16; ++u)
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = fma(d[u], xs[(k + u) * NR + q], acc[q]);
      }
      for (; k < lf.len; ++k) {
        const double dk = D[r + (int64_t)k * a.lm];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = fma(dk, xs[k * NR + q], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < NR; ++q) a.Y[lf.start + r + (int64_t)q * a.ldy] = acc[q];
    }
    gp V = (gp)a.V + lf.start;
    for (int kb = 0; kb < lf.rV; kb += KG) {
      double v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = 0.0;
      for (int t = 0; t < rpt; ++t) {
        const int r = tid + t * kMvWG;
        const bool live = r < lf.len;
#pragma unroll
        for (int kk = 0; kk < KG; ++kk) {
          const int c = kb + kk;
          const double vc = (live && c < lf.rV) ? V[(live ? r : 0) + (int64_t)c * a.N] : 0.0;
#pragma unroll
          for (int q = 0; q < NR; ++q) v[kk * NR + q] = fma(vc, live ? xs[r * NR + q] : 0.0, v[kk * NR + q]);
        }
      }
      const double sum = mv_reduce16(v);  // lane & 15 = kk * NR + q
      if (lane < 16) red[w * kMvMaxRank * NR + kb * NR + lane] = sum;
    }
    __syncthreads();
    for (int e = tid; e < lf.rV * NR; e += kMvWG) {
      double s = 0.0;
#pragma unroll
      for (int u = 0; u < kMvWaves; ++u) s += red[u * kMvMaxRank * NR + e];
      st_rel(a.G + lf.id * vst + e, s);
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
    const bool ok = T.kind == 0 ? mv_up<NR>(a, T, vs, &sh) : mv_down<NR>(a, T, vs, red2, &sh);
    if (!ok) return;
    __syncthreads();
  }

  // 3. Y += U F over the owned leaves (the first 16 columns of U of the
  // first row block loaded before the wait)
  for (int li = blockIdx.x; li < a.nleaf; li += gridDim.x) {
    const HmvLeaf lf = a.leaves[li];
    gp U = (gp)a.U + lf.start;
    const int np = (tid < lf.len) ? min(16, lf.rU) : 0;
    double up[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) up[u] = u < np ? U[tid + (int64_t)u * a.N] : 0.0;
    if (!mv_wait(a.fflag + lf.id, nullptr, nullptr, a.err, a.spin_limit, &sh)) return;
    for (int e = tid; e < lf.rU * NR; e += kMvWG) vs[e] = ld_rel(a.F + lf.id * vst + e);
    __syncthreads();
    for (int t = 0; t < rpt; ++t) {
      const int r = tid + t * kMvWG;
      if (r >= lf.len) break;
      double acc[NR];
#pragma unroll
      for (int q = 0; q < NR; ++q) acc[q] = 0.0;
      int k = 0;
      if (t == 0) {
#pragma unroll
        for (int u = 0; u < 16; ++u)
          if (u < np)
#pragma unroll
            for (int q = 0; q < NR; ++q) acc[q] = fma(up[u], vs[u * NR + q], acc[q]);
        k = np;
      }
      for (; k + 8 <= lf.rU; k += 8) {
        double e[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) e[u] = U[r + (int64_t)(k + u) * a.N];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int q = 0; q < NR; ++q) acc[q] = fma(e[u], vs[(k + u) * NR + q], acc[q]);
      }
      for (; k < lf.rU; ++k) {
        const double ek = U[r + (int64_t)k * a.N];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] = fma(ek, vs[k * NR + q], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < NR; ++q) a.Y[lf.start + r + (int64_t)q * a.ldy] += acc[q];
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

template <int NR>
static int mv_grid_for(const HmvPlan& p) {
  const size_t lds = (size_t)mv_lds_doubles(p.lm, NR) * sizeof(double);
  if (lds > 60 * 1024) return 0;
  int dev = 0, ncu = 0, coop = 0, nb = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  if (!coop) return 0;
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_hss_matvec<NR>, kMvWG, lds));
  return std::min(p.nleaf, nb * ncu);
}

int hss_mv_grid(const HmvPlan& p, int l) {
  if (!p.ok || p.lm > (int64_t)kMvWG * kMvMaxRpt || p.ldv > kMvMaxRank) return 0;
  switch (l) {
    case 0: return mv_grid_for<1>(p);
    case 1: return mv_grid_for<2>(p);
    case 2: return mv_grid_for<4>(p);
    case 3: return mv_grid_for<8>(p);
    default: return mv_grid_for<16>(p);
  }
}

template <int NR>
static hipError_t mv_launch_nr(HmvArgs a, int grid, hipStream_t s) {
  void* args[] = {&a};
  const size_t lds = (size_t)mv_lds_doubles(a.lm, NR) * sizeof(double);
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_hss_matvec<NR>), dim3(grid), dim3(kMvWG), args,
                                    (unsigned)lds, s);
}

bool hss_mv_launch(HmvPlan& p, const double* D, const double* U, const double* V, const double* pool, int64_t N,
                   const double* X, int64_t ldx, int ncols, double* Y, int64_t ldy, hipStream_t s) {
  if (!p.ok) return false;
  if (!p.err_host) {
    HIP_CHECK(hipHostMalloc((void**)&p.err_host, sizeof(int), hipHostMallocDefault));
    *p.err_host = 0;
  }
  int* sy = p.sync.as<int>();
  int done = 0;
  while (done < ncols) {
    // the widest launch that fits the remaining columns and the device
    int l = 4;
    while (l > 0 && ((1 << l) > ncols - done || p.grid[l] <= 0)) --l;
    const int nr = 1 << l, grid = p.grid[l];
    if (grid <= 0) return false;
    HmvArgs a;
    a.leaves = p.leaves.get();
    a.tasks = p.tasks.get();
    a.D = D;
    a.U = U;
    a.V = V;
    a.pool = pool;
    a.X = X + (int64_t)done * ldx;
    a.Y = Y + (int64_t)done * ldy;
    a.G = p.vec.as<double>();
    a.F = a.G + (int64_t)p.nn * p.ldv * nr;
    a.gflag = sy;
    a.fflag = sy + p.nn;
    a.counter = sy + 2 * p.nn;
    a.err = sy + 2 * p.nn + 1;
    a.N = N;
    a.lm = p.lm;
    a.spin_limit = g_mv_spin.load();
    a.ldx = ldx;
    a.ldy = ldy;
    a.nleaf = p.nleaf;
    a.ntasks = p.ntasks;
    a.ldv = p.ldv;
    a.pad_ = 0;
    // flags and counter start at zero (a kernel, not a memset); err is
    // cleared once per call, so a failed group is reported after the call
    const int nz = done == 0 ? 2 * p.nn + 2 : 2 * p.nn + 1;
    k_mv_zero<<<(unsigned)std::max(1, std::min(64, (nz + 255) / 256)), 256, 0, s>>>(sy, nz);
    HIP_CHECK(hipGetLastError());
    hipError_t e;
    switch (l) {
      case 0: e = mv_launch_nr<1>(a, grid, s); break;
      case 1: e = mv_launch_nr<2>(a, grid, s); break;
      case 2: e = mv_launch_nr<4>(a, grid, s); break;
      case 3: e = mv_launch_nr<8>(a, grid, s); break;
      default: e = mv_launch_nr<16>(a, grid, s); break;
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    done += nr;
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
