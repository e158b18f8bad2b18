// ULV factorisation and solve of the nested HSS operator on gfx950.
//
// No reference counterpart: the reference solves HSS systems by the
// push-through Woodbury recursion (include/hss_routines_omp.h:247-346),
// which stays the engine's default solver.  This is the ULV elimination of
// Chandrasekaran, Gu and Pals (SIMAX 28(3), 2006) on the engine's
// orthonormal bases: orthogonal transforms and triangular solves only, so a
// solve is backward stable without refinement.
//
// Factor, one depth per step, children before parents.  Node i holds an
// n x n block D_i (a leaf block, or the reduced block
// [[Dt_a, B_ab Vt_b^T], [B_ba Vt_a^T, Dt_b]] of its children), a row basis
// U_i (leaf basis or transfer R_i, n x r) and a column basis V_i (leaf
// basis or [Vt_a W_i(a); Vt_b W_i(b)], n x s); the root has r = s = 0.
//   U_i = H_U [R_u; 0] (Householder), Qt = diag(sign(diag R_u), I) H_U^T,
//   C = Qt D_i, whose last m = n - r rows carry no coupling;
//   C[r:]^T = H_P [Rq; 0];  G = H_P^T C[:r]^T: E = G[:m]^T, Dt = G[m:]^T;
//   H_P^T V_i = [Vh; Vt].
// Solve, up the tree: beta = b(I_i) at a leaf, [bt_a - B_ab gk_b;
// bt_b - B_ba gk_a] at a parent; gamma = Qt beta; z = Rq^{-T} gamma[r:];
// bt = gamma[:r] - E z; gk = Vh^T z + W^T [gk_a; gk_b] (the known part of
// V_i^T x).  Root: Rq^T z = beta.  Down the tree H_P [z; xt] gives the
// children's xt and, at the leaves, x.
//
// The three kernels below hold one column in the registers of one wave
// (64 lanes x PER rows, up to 1024 rows): a batched Householder QR (one
// workgroup per block; one wave forms each reflector, the four waves update
// the trailing columns), the application of a block's reflectors to
// columns, and forward substitution with Rq^T.  The small products run as
// node tasks.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hss.hpp"
#include "sweep.hpp"

namespace hmat {

using kern::BGemm;
using kern::Rect;

namespace {

constexpr int kUlvWG = 256;
constexpr int kUlvMaxRows = 1024;

struct HQr {
  double* a;      // rows x cols, column-major: R above the diagonal, reflectors below
  double* tau;    // cols
  double* sgn;    // sign of each diagonal entry of R, or null
  int* info;      // 1 + first column with a zero diagonal entry of R, or null
  int32_t rows, cols, lda, pad_;
};
// y = Q^T x (trans) or y = Q x, Q = H_0 H_1 ... H_{nref-1} (reflectors stored
// below the diagonal of v, unit diagonal implied), with the first nref rows
// of Q^T x multiplied by sgn (for trans = 0 the same rows of x, before Q).
// Rows [0, n0) of x come from x0, rows [n0, rows) from x1.  y may overlap
// x0: every lane reads its column before any write.
struct HApply {
  const double* v;
  const double* tau;
  const double* sgn;
  const double* x0;
  const double* x1;
  double* y;
  int32_t rows, nref, ldv, n0, ldx0, ldx1, ldy, ncols, trans, pad_;
};
// x = R^{-T} x in place, R upper triangular m x m
struct TrItem {
  const double* r;
  double* x;
  int32_t m, ldr, ldx, ncols;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int PER>
__global__ __launch_bounds__(kUlvWG) void k_house_qr(const HQr* __restrict__ items) {
  __shared__ double vs[PER * 64];
  __shared__ double tau_s;
  const HQr it = items[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int kmin = min(it.rows, it.cols);
  if (it.info && threadIdx.x == 0) *it.info = 0;
  for (int j = 0; j < kmin; ++j) {
    const int len = it.rows - j;
    if (w == 0) {
      // reflector of column j (LAPACK's convention: beta = -sign(alpha) ||x||)
      double* col = it.a + j + (int64_t)j * it.lda;
      double x[PER];
      double ss = 0.0;
#pragma unroll
      for (int t = 0; t < PER; ++t) {
        const int i = lane + 64 * t;
        x[t] = i < len ? col[i] : 0.0;
        if (i > 0) ss = fma(x[t], x[t], ss);
      }
      ss = wave_sum(ss);
      const double alpha = __shfl(x[0], 0, 64);
      double beta = alpha, tau = 0.0, scale = 0.0;
      if (ss > 0.0) {
        beta = -copysign(sqrt(fma(alpha, alpha, ss)), alpha);
        tau = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
#pragma unroll
      for (int t = 0; t < PER; ++t) {
        const int i = lane + 64 * t;
        if (i < len) {
          const double v = i == 0 ? 1.0 : x[t] * scale;
          vs[i] = v;
          col[i] = i == 0 ? beta : v;
        }
      }
      if (lane == 0) {
        it.tau[j] = tau;
        if (it.sgn) it.sgn[j] = beta < 0.0 ? -1.0 : 1.0;
        if (it.info && beta == 0.0 && *it.info == 0) *it.info = j + 1;
        tau_s = tau;
      }
    }
    __syncthreads();
    const double tau = tau_s;
    if (tau != 0.0)
      for (int k = j + 1 + w; k < it.cols; k += kUlvWG / 64) {
        double* c = it.a + j + (int64_t)k * it.lda;
        double cv[PER];
        double dot = 0.0;
#pragma unroll
        for (int t = 0; t < PER; ++t) {
          const int i = lane + 64 * t;
          cv[t] = i < len ? c[i] : 0.0;
          if (i < len) dot = fma(vs[i], cv[t], dot);
        }
        dot = wave_sum(dot) * tau;
#pragma unroll
        for (int t = 0; t < PER; ++t) {
          const int i = lane + 64 * t;
          if (i < len) c[i] = fma(-dot, vs[i], cv[t]);
        }
      }
    __syncthreads();  // column j + 1 and vs are read next
  }
}

template <int PER>
__device__ __forceinline__ void load_reflector(const HApply& it, int j, int lane, double (&v)[PER]) {
  const double* vj = it.v + (int64_t)j * it.ldv;
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int i = lane + 64 * t;
    v[t] = (i > j && i < it.rows) ? vj[i] : (i == j ? 1.0 : 0.0);
  }
}

template <int PER>
__global__ __launch_bounds__(kUlvWG) void k_house_apply(const HApply* __restrict__ items) {
  const HApply it = items[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * (kUlvWG / 64) + (threadIdx.x >> 6);
  if (col >= it.ncols) return;  // no barrier below
  double c[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int i = lane + 64 * t;
    c[t] = i < it.n0 ? it.x0[i + (int64_t)col * it.ldx0]
                     : (i < it.rows ? it.x1[(i - it.n0) + (int64_t)col * it.ldx1] : 0.0);
  }
  if (it.sgn && !it.trans)
#pragma unroll
    for (int t = 0; t < PER; ++t)
      if (lane + 64 * t < it.nref) c[t] *= it.sgn[lane + 64 * t];
  if (it.nref > 0) {
    // reflector q + 1 is loaded while q is applied
    double v[PER], vn[PER];
    int j = it.trans ? 0 : it.nref - 1;
    load_reflector<PER>(it, j, lane, v);
    double tj = it.tau[j];
    for (int q = 0; q < it.nref; ++q) {
      const int jn = it.trans ? j + 1 : j - 1;
      const bool more = q + 1 < it.nref;
      double tn = 0.0;
      if (more) {
        load_reflector<PER>(it, jn, lane, vn);
        tn = it.tau[jn];
      }
      double dot = 0.0;
#pragma unroll
      for (int t = 0; t < PER; ++t) dot = fma(v[t], c[t], dot);
      dot = wave_sum(dot) * tj;
#pragma unroll
      for (int t = 0; t < PER; ++t) c[t] = fma(-dot, v[t], c[t]);
      if (more) {
#pragma unroll
        for (int t = 0; t < PER; ++t) v[t] = vn[t];
        tj = tn;
        j = jn;
      }
    }
  }
  if (it.sgn && it.trans)
#pragma unroll
    for (int t = 0; t < PER; ++t)
      if (lane + 64 * t < it.nref) c[t] *= it.sgn[lane + 64 * t];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int i = lane + 64 * t;
    if (i < it.rows) it.y[i + (int64_t)col * it.ldy] = c[t];
  }
}

// z_k = (x_k - sum_{j<k} R(j, k) z_j) / R(k, k): column k of R is
// contiguous; column k + 1 is loaded while step k reduces.
template <int PER>
__global__ __launch_bounds__(kUlvWG) void k_ulv_trsv(const TrItem* __restrict__ items) {
  const TrItem it = items[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * (kUlvWG / 64) + (threadIdx.x >> 6);
  if (col >= it.ncols || it.m == 0) return;  // no barrier below
  double* x = it.x + (int64_t)col * it.ldx;
  double xb[PER], z[PER], rc[PER], rn[PER];
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int i = lane + 64 * t;
    xb[t] = i < it.m ? x[i] : 0.0;
    z[t] = 0.0;
    rc[t] = 0.0;  // column 0 has no entries above its diagonal
  }
  double dk = it.r[0];
  for (int k = 0; k < it.m; ++k) {
    const bool more = k + 1 < it.m;
    double dn = 1.0;
    if (more) {
      const double* r1 = it.r + (int64_t)(k + 1) * it.ldr;
#pragma unroll
      for (int t = 0; t < PER; ++t) {
        const int i = lane + 64 * t;
        rn[t] = i <= k ? r1[i] : 0.0;
      }
      dn = r1[k + 1];
    }
    double s = 0.0;
#pragma unroll
    for (int t = 0; t < PER; ++t) s = fma(rc[t], z[t], s);
    s = wave_sum(s);
    const int tk = k >> 6, lk = k & 63;
    double xk = 0.0;
#pragma unroll
    for (int t = 0; t < PER; ++t)
      if (t == tk) xk = xb[t];
    xk = __shfl(xk, lk, 64);
    const double zk = (xk - s) / dk;
#pragma unroll
    for (int t = 0; t < PER; ++t)
      if (t == tk && lane == lk) z[t] = zk;
    if (more) {
#pragma unroll
      for (int t = 0; t < PER; ++t) rc[t] = rn[t];
      dk = dn;
    }
  }
#pragma unroll
  for (int t = 0; t < PER; ++t) {
    const int i = lane + 64 * t;
    if (i < it.m) x[i] = z[t];
  }
}

int rows_per_lane(int rows) { return rows <= 256 ? 4 : (rows <= 512 ? 8 : 16); }

void launch_qr(const std::vector<HQr>& v, hipStream_t s) {
  if (v.empty()) return;
  int rmax = 0;
  for (const auto& x : v) rmax = std::max(rmax, (int)x.rows);
  const HQr* d = stage(v, s);
  const unsigned g = (unsigned)v.size();
  switch (rows_per_lane(rmax)) {
    case 4: k_house_qr<4><<<g, kUlvWG, 0, s>>>(d); break;
    case 8: k_house_qr<8><<<g, kUlvWG, 0, s>>>(d); break;
    default: k_house_qr<16><<<g, kUlvWG, 0, s>>>(d); break;
  }
  HIP_CHECK(hipGetLastError());
}

// Reflector applications of `count` descriptors already on the device, up to
// rmax rows and cmax columns each.
void launch_apply_dev(const HApply* d, int count, int rmax, int cmax, hipStream_t s) {
  if (count == 0 || cmax == 0) return;
  const dim3 g((unsigned)((cmax + kUlvWG / 64 - 1) / (kUlvWG / 64)), (unsigned)count);
  switch (rows_per_lane(rmax)) {
    case 4: k_house_apply<4><<<g, kUlvWG, 0, s>>>(d); break;
    case 8: k_house_apply<8><<<g, kUlvWG, 0, s>>>(d); break;
    default: k_house_apply<16><<<g, kUlvWG, 0, s>>>(d); break;
  }
  HIP_CHECK(hipGetLastError());
}

void apply_extent(const std::vector<HApply>& v, int& rmax, int& cmax) {
  rmax = cmax = 0;
  for (const auto& x : v) {
    rmax = std::max(rmax, (int)x.rows);
    cmax = std::max(cmax, (int)x.ncols);
  }
}

void launch_apply(const std::vector<HApply>& v, hipStream_t s) {
  int rmax, cmax;
  apply_extent(v, rmax, cmax);
  if (v.empty() || cmax == 0) return;
  launch_apply_dev(stage(v, s), (int)v.size(), rmax, cmax, s);
}

// Forward substitutions of `count` descriptors already on the device.
void launch_trsv_dev(const TrItem* d, int count, int mmax, int cmax, hipStream_t s) {
  if (count == 0) return;
  const dim3 g((unsigned)((cmax + kUlvWG / 64 - 1) / (kUlvWG / 64)), (unsigned)count);
  switch (rows_per_lane(mmax)) {
    case 4: k_ulv_trsv<4><<<g, kUlvWG, 0, s>>>(d); break;
    case 8: k_ulv_trsv<8><<<g, kUlvWG, 0, s>>>(d); break;
    default: k_ulv_trsv<16><<<g, kUlvWG, 0, s>>>(d); break;
  }
  HIP_CHECK(hipGetLastError());
}

int64_t align32(int64_t k) { return (std::max<int64_t>(k, 1) + 31) / 32 * 32; }

}  // namespace

void HssEngine::ulv_factor() {
  const int nn = (int)tree.nodes.size();
  const int64_t lm = tree.lmax;
  un_.assign(nn, UNode{});
  int64_t tot = 0, cmax = 0;
  auto take = [&](int64_t k) {
    const int64_t o = tot;
    tot += align32(k);
    return o;
  };
  for (int d = tree.max_depth; d >= 0; --d) {
    int64_t cs = 0;
    for (int id : tree.by_depth[d]) {
      const TNode& nd = tree.nodes[id];
      UNode& u = un_[id];
      u.n = nd.leaf ? (int)nd.len : un_[nd.left].r + un_[nd.right].r;
      u.r = id == 0 ? 0 : sn[id].rU;
      u.s = id == 0 ? 0 : sn[id].rV;
      HMAT_REQUIRE(u.n <= kUlvMaxRows, HMAT_ERR_ARG,
                   "ULV: a block of " + std::to_string(u.n) + " rows at depth " + std::to_string(d) +
                       " exceeds " + std::to_string(kUlvMaxRows) +
                       " (leaves above 1024 points, or ranks that add up as in the reference-mode build)");
      HMAT_REQUIRE(u.r <= u.n && (id == 0 || (u.r >= 1 && u.s >= 1)), HMAT_ERR_STATE,
                   "ULV: node rank outside [1, block rows]");
      u.m = u.n - u.r;
      u.oHU = take((int64_t)u.n * u.r);
      u.oTU = take(u.r);
      u.oSU = take(u.r);
      u.oHP = take((int64_t)u.n * u.m);
      u.oTP = take(u.m);
      u.oG = take((int64_t)u.n * u.r);
      u.oV = take((int64_t)u.n * u.s);
      u.oDt = take((int64_t)u.r * u.r);
      cs += align32((int64_t)u.n * u.n);
    }
    cmax = std::max(cmax, cs);
  }
  upool_.reserve((size_t)tot);
  uC_.reserve((size_t)cmax);
  uinfo_.reserve((size_t)nn);
  zero_bytes_async(uinfo_.get(), (size_t)nn * sizeof(int), stream);
  double* UP = upool_.get();
  const double* P = pool.get();
  double* Cb = uC_.get();
  for (int d = tree.max_depth; d >= 0; --d) {
    const auto& ids = tree.by_depth[d];
    std::vector<int64_t> oC(ids.size());
    int64_t cs = 0;
    for (size_t q = 0; q < ids.size(); ++q) {
      oC[q] = cs;
      cs += align32((int64_t)un_[ids[q]].n * un_[ids[q]].n);
    }
    // the blocks, bases and column bases of this depth
    SweepProgram sw;
    sw.begin_phase();
    for (size_t q = 0; q < ids.size(); ++q) {
      const int id = ids[q];
      const TNode& nd = tree.nodes[id];
      const UNode& u = un_[id];
      double* C = Cb + oC[q];
      sw.begin_task();
      if (nd.leaf) {
        sw.copy(D.get() + (int64_t)nd.leaf_idx * lm * lm, (int)lm, C, u.n, u.n, u.n);
        if (id != 0) {
          sw.copy(Ul.get() + nd.start, (int)N, UP + u.oHU, u.n, u.n, u.r);
          sw.copy(Vl.get() + nd.start, (int)N, UP + u.oV, u.n, u.n, u.s);
        }
      } else {
        const UNode &ua = un_[nd.left], &ub = un_[nd.right];
        sw.copy(UP + ua.oDt, ua.r, C, u.n, ua.r, ua.r);
        sw.copy(UP + ub.oDt, ub.r, C + ua.r + (int64_t)ua.r * u.n, u.n, ub.r, ub.r);
        // C(a, b) = B_ab Vt_b^T, C(b, a) = B_ba Vt_a^T
        sw.op(BGemm{P + sn[id].oBab, UP + ub.oV + ub.m, C + (int64_t)ua.r * u.n, ua.r, ub.r, ub.s, sn[id].ldB, ub.n,
                    u.n, 0, 1, 1.0, 0.0});
        sw.op(BGemm{P + sn[id].oBba, UP + ua.oV + ua.m, C + ua.r, ub.r, ua.r, ua.s, sn[id].ldB, ua.n, u.n, 0, 1,
                    1.0, 0.0});
        if (id != 0) {
          sw.copy(P + sn[id].oR, sn[id].ldR, UP + u.oHU, u.n, u.n, u.r);
          sw.op(BGemm{UP + ua.oV + ua.m, P + sn[id].oW, UP + u.oV, ua.r, u.s, ua.s, ua.n, sn[id].ldW, u.n, 0, 0, 1.0,
                      0.0});
          sw.op(BGemm{UP + ub.oV + ub.m, P + sn[id].oW + ua.s, UP + u.oV + ua.r, ub.r, u.s, ub.s, ub.n, sn[id].ldW,
                      u.n, 0, 0, 1.0, 0.0});
        }
      }
      sw.end_task();
    }
    sw.run(stream);
    std::vector<HQr> qu, qp;
    std::vector<HApply> au, ap;
    std::vector<Rect> tr, tdt;
    for (size_t q = 0; q < ids.size(); ++q) {
      const int id = ids[q];
      const UNode& u = un_[id];
      double* C = Cb + oC[q];
      if (u.r > 0) {
        qu.push_back(HQr{UP + u.oHU, UP + u.oTU, UP + u.oSU, nullptr, u.n, u.r, u.n, 0});
        au.push_back(HApply{UP + u.oHU, UP + u.oTU, UP + u.oSU, C, nullptr, C, u.n, u.r, u.n, u.n, u.n, 0, u.n, u.n, 1,
                            0});
        tr.push_back(Rect{C, UP + u.oG, u.n, u.n, u.r, u.n, 1.0, 0.0});
        tdt.push_back(Rect{UP + u.oG + u.m, UP + u.oDt, u.n, u.r, u.r, u.r, 1.0, 0.0});
      }
      if (u.m > 0) {
        tr.push_back(Rect{C + u.r, UP + u.oHP, u.n, u.n, u.m, u.n, 1.0, 0.0});
        qp.push_back(HQr{UP + u.oHP, UP + u.oTP, nullptr, uinfo_.get() + id, u.n, u.m, u.n, 0});
        if (u.r > 0)
          ap.push_back(HApply{UP + u.oHP, UP + u.oTP, nullptr, UP + u.oG, nullptr, UP + u.oG, u.n, u.m, u.n, u.n, u.n,
                              0, u.n, u.r, 1, 0});
        if (u.s > 0)
          ap.push_back(HApply{UP + u.oHP, UP + u.oTP, nullptr, UP + u.oV, nullptr, UP + u.oV, u.n, u.m, u.n, u.n, u.n,
                              0, u.n, u.s, 1, 0});
      }
    }
    launch_qr(qu, stream);              // U_i = H_U [R_u; 0]
    launch_apply(au, stream);           // C = Qt D_i
    kern::rect_transpose(tr, drc_, stream);  // C[r:]^T, C[:r]^T
    launch_qr(qp, stream);              // C[r:]^T = H_P [Rq; 0]
    launch_apply(ap, stream);           // G = H_P^T C[:r]^T, H_P^T V_i
    kern::rect_transpose(tdt, drc_, stream);  // Dt = G[m:]^T
  }
  auto info = download(uinfo_.get(), (size_t)nn, stream);
  for (int id = 0; id < nn; ++id)
    if (info[id] != 0)
      throw Error(HMAT_ERR_SINGULAR, "ULV: singular reduced block at depth " + std::to_string(tree.nodes[id].depth) +
                                         " (zero pivot in column " + std::to_string(info[id] - 1) + ")");
  uzero_.reserve((size_t)std::max(rVmax, 1));
  zero_bytes_async(uzero_.get(), (size_t)std::max(rVmax, 1) * sizeof(double), stream);
}

double HssEngine::ulv_bytes(int nrhs) const {
  double f = 0.0;
  for (size_t id = 0; id < un_.size(); ++id) {
    const UNode& u = un_[id];
    f += (double)u.n * u.r + (double)u.n * u.m + (double)u.m * u.r + (double)u.m * u.s;
    const TNode& nd = tree.nodes[id];
    if (!nd.leaf) {
      const SNode &a = sn[nd.left], &b = sn[nd.right];
      f += (double)a.rU * b.rV + (double)b.rU * a.rV;
      if (id != 0) f += (double)(a.rV + b.rV) * sn[id].rV;
    }
  }
  return 8.0 * f + 16.0 * (double)N * nrhs;
}

void HssEngine::ulv_apply(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx) {
  ProfScope prof("hss_ulv_apply", stream, ulv_bytes(nrhs));
  const int nn = (int)tree.nodes.size();
  std::vector<int64_t> ob(nn), og(nn);
  int64_t tb = 0, tg = 0;
  for (int id = 0; id < nn; ++id) {
    ob[id] = tb;
    tb += align32((int64_t)un_[id].n * nrhs);
    og[id] = tg;
    tg += align32((int64_t)std::max(un_[id].s, 1) * nrhs);
  }
  ubeta_.reserve((size_t)tb);
  ugk_.reserve((size_t)tg);
  double* UP = upool_.get();
  const double* P = pool.get();
  // the launches' descriptors are built on the first call for these
  // buffers and column count, and replayed afterwards
  const std::vector<const void*> key = {dB, dX, reinterpret_cast<const void*>(static_cast<intptr_t>(ldb)),
                                        reinterpret_cast<const void*>(static_cast<intptr_t>(ldx)), UP, P,
                                        ubeta_.get(), ugk_.get(), uzero_.get()};
  if (!ulvp_.ok || ulvp_.nrhs != nrhs || ulvp_.key != key) {
    ulvp_.ok = false;
    auto beta = [&](int id) { return ubeta_.get() + ob[id]; };
    auto gk = [&](int id) { return ugk_.get() + og[id]; };
    std::vector<HApply> ha;
    std::vector<TrItem> tr;
    std::vector<BGemm> ops;
    std::vector<kern::NodeTask> tasks;
    std::vector<UlvStep> steps;
    auto add_apply = [&](const std::vector<HApply>& v) {
      int rmax, cmax;
      apply_extent(v, rmax, cmax);
      if (v.empty() || cmax == 0) return;
      steps.push_back(UlvStep{0, (int32_t)ha.size(), (int32_t)v.size(), rmax, cmax});
      ha.insert(ha.end(), v.begin(), v.end());
    };
    // up the tree
    for (int d = tree.max_depth; d >= 0; --d) {
      std::vector<HApply> au;
      std::vector<TrItem> ts;
      for (int id : tree.by_depth[d]) {
        const TNode& nd = tree.nodes[id];
        const UNode& u = un_[id];
        if (nd.leaf)
          au.push_back(HApply{UP + u.oHU, UP + u.oTU, UP + u.oSU, dB + nd.start, nullptr, beta(id), u.n, u.r, u.n,
                              u.n, (int)ldb, 0, u.n, nrhs, 1, 0});
        else if (u.r > 0)
          au.push_back(HApply{UP + u.oHU, UP + u.oTU, UP + u.oSU, beta(id), nullptr, beta(id), u.n, u.r, u.n, u.n,
                              u.n, 0, u.n, nrhs, 1, 0});
        if (u.m > 0) ts.push_back(TrItem{UP + u.oHP, beta(id) + u.r, u.m, u.n, u.n, nrhs});
      }
      add_apply(au);  // gamma = Qt beta
      if (!ts.empty()) {  // z = Rq^{-T} gamma[r:]
        int mmax = 0, cmax = 0;
        for (const auto& x : ts) {
          mmax = std::max(mmax, (int)x.m);
          cmax = std::max(cmax, (int)x.ncols);
        }
        steps.push_back(UlvStep{1, (int32_t)tr.size(), (int32_t)ts.size(), mmax, cmax});
        tr.insert(tr.end(), ts.begin(), ts.end());
      }
      if (d == 0) break;
      // one task per parent: finish both children (bt, gk), then the
      // parent's right-hand side
      SweepProgram sw;
      sw.begin_phase();
      for (int p : tree.by_depth[d - 1]) {
        const TNode& np = tree.nodes[p];
        if (np.leaf) continue;
        sw.begin_task();
        for (int c : {np.left, np.right}) {
          const TNode& nc = tree.nodes[c];
          const UNode& u = un_[c];
          if (u.m > 0)  // bt = gamma[:r] - E z
            sw.op(BGemm{UP + u.oG, beta(c) + u.r, beta(c), u.r, nrhs, u.m, u.n, u.n, u.n, 1, 0, -1.0, 1.0});
          bool first = true;
          if (!nc.leaf) {  // gk = W^T [gk_a; gk_b] + Vh^T z
            const UNode &ua = un_[nc.left], &ub = un_[nc.right];
            sw.op(BGemm{P + sn[c].oW, gk(nc.left), gk(c), u.s, nrhs, ua.s, sn[c].ldW, ua.s, u.s, 1, 0, 1.0, 0.0});
            sw.op(BGemm{P + sn[c].oW + ua.s, gk(nc.right), gk(c), u.s, nrhs, ub.s, sn[c].ldW, ub.s, u.s, 1, 0, 1.0,
                        1.0});
            first = false;
          }
          if (u.m > 0) {
            sw.op(BGemm{UP + u.oV, beta(c) + u.r, gk(c), u.s, nrhs, u.m, u.n, u.n, u.s, 1, 0, 1.0,
                        first ? 0.0 : 1.0});
            first = false;
          }
          if (first) sw.copy(uzero_.get(), 0, gk(c), u.s, u.s, nrhs);  // nothing known below a full-rank leaf
        }
        const UNode &ua = un_[np.left], &ub = un_[np.right], &up = un_[p];
        double* bp = beta(p);
        sw.copy(beta(np.left), ua.n, bp, up.n, ua.r, nrhs);
        sw.op(BGemm{P + sn[p].oBab, gk(np.right), bp, ua.r, nrhs, ub.s, sn[p].ldB, ub.s, up.n, 0, 0, -1.0, 1.0});
        sw.copy(beta(np.right), ub.n, bp + ua.r, up.n, ub.r, nrhs);
        sw.op(BGemm{P + sn[p].oBba, gk(np.left), bp + ua.r, ub.r, nrhs, ua.s, sn[p].ldB, ua.s, up.n, 0, 0, -1.0, 1.0});
        sw.end_task();
      }
      if (!sw.tasks.empty()) {
        const int op0 = (int)ops.size();
        steps.push_back(UlvStep{2, (int32_t)tasks.size(), (int32_t)sw.tasks.size(), 0, 0});
        for (kern::NodeTask t : sw.tasks) {
          t.op0 += op0;
          tasks.push_back(t);
        }
        ops.insert(ops.end(), sw.ops.begin(), sw.ops.end());
      }
    }
    // down the tree: y = H_P [z; xt]
    for (int d = 0; d <= tree.max_depth; ++d) {
      std::vector<HApply> ad;
      for (int id : tree.by_depth[d]) {
        const TNode& nd = tree.nodes[id];
        const UNode& u = un_[id];
        const double* xt = nullptr;
        int ldxt = 0;
        if (id != 0) {
          const TNode& npar = tree.nodes[nd.parent];
          xt = beta(nd.parent) + (npar.right == id ? un_[npar.left].r : 0);
          ldxt = un_[nd.parent].n;
        }
        double* y = nd.leaf ? dX + nd.start : beta(id);
        const int ldy = nd.leaf ? (int)ldx : u.n;
        ad.push_back(HApply{UP + u.oHP, UP + u.oTP, nullptr, beta(id) + u.r, xt, y, u.n, u.m, u.n, u.m, u.n, ldxt,
                            ldy, nrhs, 0, 0});
      }
      add_apply(ad);
    }
    const size_t hab = ha.size() * sizeof(HApply);
    ulvp_.tr_off = (int64_t)((hab + 255) / 256 * 256);
    std::vector<unsigned char> bytes((size_t)ulvp_.tr_off + tr.size() * sizeof(TrItem));
    if (!ha.empty()) std::memcpy(bytes.data(), ha.data(), hab);
    if (!tr.empty()) std::memcpy(bytes.data() + ulvp_.tr_off, tr.data(), tr.size() * sizeof(TrItem));
    ulvp_.items.upload(bytes, stream);
    if (!ops.empty()) ulvp_.ops.upload(ops, stream);
    if (!tasks.empty()) ulvp_.tasks.upload(tasks, stream);
    ulvp_.steps = steps;
    ulvp_.nrhs = nrhs;
    ulvp_.key = key;
    ulvp_.ok = true;
  }
  const HApply* dha = reinterpret_cast<const HApply*>(ulvp_.items.get());
  const TrItem* dtr = reinterpret_cast<const TrItem*>(ulvp_.items.get() + ulvp_.tr_off);
  for (const UlvStep& st : ulvp_.steps) {
    if (st.kind == 0)
      launch_apply_dev(dha + st.off, st.count, st.rows, st.cols, stream);
    else if (st.kind == 1)
      launch_trsv_dev(dtr + st.off, st.count, st.rows, st.cols, stream);
    else
      kern::node_tasks(ulvp_.tasks.get() + st.off, st.count, ulvp_.ops.get(), stream);
  }
  prof.finish();
  HIP_CHECK(hipStreamSynchronize(stream));
}

}  // namespace hmat
