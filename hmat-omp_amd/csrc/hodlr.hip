// HODLR engine on gfx950.
//
// Reference: HODLRMatrixOMP (include/hodlr_routines_omp.h).  The reference
// recurses node by node (one OpenMP task per subtree, one dgemm per product)
// and refactorises every leaf on every Woodbury visit (4^L leaf LUs).  Here:
//   * build: every off-diagonal block of a tree depth is compressed in one
//     batch (generated sketch on MFMA, CholeskyQR3, Jacobi SVD of k x k);
//   * matvec: level-batched V^T x then U S t;
//   * solve: the Woodbury identity of hodlr_routines_omp.h:147-240 in its
//     push-through form (I + K C), factored once bottom-up, then applied as
//     a sweep from the deepest level to the root.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <type_traits>

#include "hodlr.hpp"

namespace hmat {

HodlrEngine::HodlrEngine(const hmat_options& o) : opt(o) {
  HIP_CHECK(hipGetDevice(&device));
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  own_stream = true;
  if (opt.precision == HMAT_PREC_MIXED) refine_rtol = 1e-10;
}

HodlrEngine::~HodlrEngine() {
  if (stream) stager_release(stream);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

// Post-order over internal nodes, A12 (rows = left child) then A21 (rows =
// right child), only blocks that take the randomised branch
// (hodlr_routines.h:43-53).  Returns (omega offset, k) per block in order,
// keyed by the row node id through the parallel vector `rows`.
std::vector<std::pair<int64_t, int>> HodlrEngine::omega_schedule(const hmat_options& o, int64_t N,
                                                                 int leaf) {
  std::vector<std::pair<int64_t, int>> out;  // (n, k) in draw order
  struct Rec {
    const hmat_options& o;
    int64_t leaf;
    std::vector<std::pair<int64_t, int>>& out;
    void walk(int64_t s, int64_t e) {
      if (e - s <= leaf) return;
      const int64_t mid = (s + e) / 2;
      walk(s, mid);
      walk(mid, e);
      const int64_t ml = mid - s, mr = e - mid;
      if (uses_rsvd(ml, mr, o)) out.push_back({mr, sketch_width(ml, mr, o)});
      if (uses_rsvd(mr, ml, o)) out.push_back({ml, sketch_width(mr, ml, o)});
    }
  } rec{o, leaf, out};
  rec.walk(0, N);
  return out;
}

void HodlrEngine::reset_for_build(int leaf_size) {
  HMAT_REQUIRE(leaf_size > 0, HMAT_ERR_ARG, "leaf_size must be positive");
  if (!fixed_) bg_.drop();  // a recorded batch graph writes this handle's arrays
  HMAT_REQUIRE(opt.rank > 0 && opt.oversample >= 0 && opt.power_iters >= 0, HMAT_ERR_ARG,
               "bad rank/oversample/power_iters");
  built = factored = false;
  f64_.clear();
  f32_.clear();
  f64T_.clear();
  f32T_.clear();
  factoredT_ = transposed_ = false;
  leaf = leaf_size;
}

void HodlrEngine::build(const GenDesc& g, int64_t n, int leaf_size, const double* host_omega,
                        int64_t omega_len) {
  HMAT_REQUIRE(n > 0, HMAT_ERR_ARG, "N must be positive");
  reset_for_build(leaf_size);
  N = n;
  tree.build(N, leaf);
  build_levels(g, host_omega, omega_len);
}

void HodlrEngine::build_forest(const GenDesc& g, const std::vector<int64_t>& sizes, int leaf_size) {
  HMAT_REQUIRE(!sizes.empty(), HMAT_ERR_ARG, "empty batch");
  int64_t tot = 0;
  for (int64_t n : sizes) {
    HMAT_REQUIRE(n > 0, HMAT_ERR_ARG, "every system size must be positive");
    tot += n;
  }
  reset_for_build(leaf_size);
  N = tot;
  tree.build_forest(sizes, leaf);
  build_levels(g, nullptr, 0);
}

void HodlrEngine::build_levels(const GenDesc& g, const double* host_omega, int64_t omega_len) {
  // levels keep their device arrays from the previous build (reallocated
  // only when a shape changes)
  lv.resize(tree.max_depth + 1);
  par.clear();

  // Omega offsets per (row node id) in the serial draw order
  std::vector<int64_t> omega_off(tree.nodes.size(), -1);
  if (host_omega) {
    int64_t pos = 0;
    struct Rec {
      HodlrEngine* e;
      std::vector<int64_t>& off;
      int64_t& pos;
      void walk(int id) {
        const TNode& nd = e->tree.nodes[id];
        if (nd.leaf) return;
        walk(nd.left);
        walk(nd.right);
        const TNode& l = e->tree.nodes[nd.left];
        const TNode& r = e->tree.nodes[nd.right];
        if (uses_rsvd(l.len, r.len, e->opt)) {
          off[nd.left] = pos;
          pos += r.len * sketch_width(l.len, r.len, e->opt);
        }
        if (uses_rsvd(r.len, l.len, e->opt)) {
          off[nd.right] = pos;
          pos += l.len * sketch_width(r.len, l.len, e->opt);
        }
      }
    } rec{this, omega_off, pos};
    rec.walk(0);
    HMAT_REQUIRE(pos == omega_len, HMAT_ERR_ARG,
                 "omega_len " + std::to_string(omega_len) + " != expected " + std::to_string(pos));
  }

  // Depths are built together in groups: every launch of a group covers
  // all its blocks.  A group's scratch (three N x k arrays per depth) is
  // kept within group_bytes (at least one depth per group), and its BF16
  // planes below 2^30 entries.
  {
    int c0 = 1;
    while (c0 <= tree.max_depth) {
      int c1 = c0;
      int kc = group_kcap(c0, c0);
      while (c1 + 1 <= tree.max_depth) {
        const int kn = std::max(kc, group_kcap(c1 + 1, c1 + 1));
        const double per = (double)N * kn * (c1 + 2 - c0);
        if (per * 24.0 > group_bytes || per * 1.0 >= (double)(1 << 29)) break;
        kc = kn;
        ++c1;
      }
      build_group(c0, c1, g, omega_off, host_omega);
      c0 = c1 + 1;
    }
  }

  // leaf diagonal blocks D (hodlr_routines_omp.h:38-40)
  const int64_t lm = tree.lmax;
  D.resize((size_t)tree.leaves.size() * lm * lm);
  D.zero(stream);
  std::vector<BlockDesc> bd;
  for (size_t i = 0; i < tree.leaves.size(); ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    bd.push_back({nd.start, nd.start, (int64_t)i * lm * lm, (int32_t)nd.len, (int32_t)nd.len, (int32_t)lm, 0});
  }
  kern::gen_blocks(g, bd, dbd_, D.get(), stream);
  if (diag_shift_) {
    std::vector<int> lens;
    for (int id : tree.leaves) lens.push_back((int)tree.nodes[id].len);
    kern::add_diag(D.get(), lm, stage(lens, stream), (int)lens.size(), diag_shift_, stream);
  }
  if (!capturing_) HIP_CHECK(hipStreamSynchronize(stream));
  // the build scratch (three N x k arrays) is kept for the next build
  prepare_matvec();
  built = true;
}

// Per-(leaf, depth) descriptors of the fused matvec: the depth-c segment
// holding each leaf, its ranks and the global index of its sibling.
void HodlrEngine::prepare_matvec() {
  const int L = tree.max_depth;
  const int nl = (int)tree.leaves.size();
  std::vector<kern::MvLeaf> lf(nl);
  std::vector<int64_t> starts(nl);
  double bytes = 0.0;
  for (int i = 0; i < nl; ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    lf[i] = kern::MvLeaf{nd.start, (int32_t)nd.len, 0};
    starts[i] = nd.start;
    bytes += 8.0 * (double)nd.len * (double)nd.len;
  }
  int KM = 1;
  for (int c = 1; c <= L; ++c) KM = std::max(KM, lv[c].rmax);
  std::vector<kern::MvLevel> lvl(L);
  std::vector<kern::MvEntry> ent((size_t)nl * L, kern::MvEntry{0, 0, 0, 0});
  std::vector<kern::MvSeg> segs;
  for (int c = 1; c <= L; ++c) {
    const HLevel& Lc = lv[c];
    lvl[c - 1] = kern::MvLevel{Lc.V.get(), Lc.US.get(), N};
    const int base = (int)segs.size();
    const int ns = (int)Lc.ids.size();
    for (int s = 0; s < ns; ++s) {
      const TNode& nd = tree.nodes[Lc.ids[s]];
      const int l0 = (int)(std::lower_bound(starts.begin(), starts.end(), nd.start) - starts.begin());
      const int l1 = (int)(std::lower_bound(starts.begin(), starts.end(), nd.start + nd.len) - starts.begin());
      const int vcols = Lc.rank[Lc.sib[s]], ucols = Lc.rank[s];
      segs.push_back(kern::MvSeg{c - 1, l0, l1, vcols});
      for (int leaf = l0; leaf < l1; ++leaf)
        ent[(size_t)leaf * L + (c - 1)] = kern::MvEntry{vcols, ucols, base + Lc.sib[s], 0};
      bytes += 8.0 * (double)nd.len * (double)(vcols + ucols);
    }
  }
  mv_leaves_.upload(lf, stream);
  if (L > 0) {
    mv_lvls_.upload(lvl, stream);
    mv_ent_.upload(ent, stream);
    mv_segs_.upload(segs, stream);
  }
  mv_.leaves = mv_leaves_.get();
  mv_.lvls = mv_lvls_.get();
  mv_.ent = mv_ent_.get();
  mv_.segs = mv_segs_.get();
  mv_.nleaf = nl;
  mv_.L = L;
  mv_.nseg = (int)segs.size();
  mv_.KM = KM;
  mv_bytes_base_ = bytes;
  // widest fused pass whose leaf-expand staging fits 64 KB of LDS
  mv_width_ = 0;
  for (int w : {16, 8, 4})
    if (kern::mv_lds_bytes(mv_, tree.lmax, w) <= 64 * 1024) {
      mv_width_ = w;
      break;
    }
  const size_t wr = (size_t)std::max(mv_width_, 4);
  mv_part_.reserve(std::max<size_t>((size_t)nl * L * KM * wr, 1));
  mv_t_.reserve(std::max<size_t>(segs.size() * KM * wr, 1));
}

double HodlrEngine::matvec_bytes(int nrhs) const { return mv_bytes_base_ + 16.0 * (double)N * nrhs; }

// Fused path, passes of 16, 8, 4, 2 or 1 right-hand sides (at most
// mv_width_ at a time): every pass reads the generators once.
void HodlrEngine::matvec_async(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy) {
  const int64_t lm = tree.lmax;
  int done = 0;
  while (done < nrhs) {
    const int rem = nrhs - done;
    int nr = 1;
    while (nr * 2 <= std::min(rem, mv_width_)) nr *= 2;
    kern::hodlr_matvec_fused(mv_, D.get(), lm, dX + (int64_t)done * ldx, ldx, nr, mv_part_.get(), mv_t_.get(),
                             dY + (int64_t)done * ldy, ldy, stream);
    done += nr;
  }
}

// Largest sketch width of the blocks of depths c0..c1.
int HodlrEngine::group_kcap(int c0, int c1) const {
  int kc = 0;
  for (int c = c0; c <= c1; ++c)
    for (int id : tree.by_depth[c]) {
      const TNode& nd = tree.nodes[id];
      const TNode& p = tree.nodes[nd.parent];
      const TNode& sb = tree.nodes[(p.left == id) ? p.right : p.left];
      const int64_t m = nd.len, n = sb.len;
      kc = std::max(kc, uses_rsvd(m, n, opt) ? sketch_width(m, n, opt) : (int)std::min(m, n));
    }
  return kc;
}

void HodlrEngine::build_group(int c0, int c1, const GenDesc& g, const std::vector<int64_t>& omega_off,
                              const double* host_omega) {
  // the segments of depths c0..c1, depth by depth; segment t of the group
  // uses column block (lvl[t] - c0) of the N x (kcap * ncl) scratch arrays
  const int ncl = c1 - c0 + 1;
  std::vector<BlockMode> mode;
  std::vector<int> k, lvl, pos;
  std::vector<int64_t> st, ln, sst, sln;
  std::vector<int> base(ncl + 1, 0);
  for (int c = c0; c <= c1; ++c) {
    HLevel& L = lv[c];
    L.ids = tree.by_depth[c];
    const int ns = (int)L.ids.size();
    L.sib.assign(ns, 0);
    base[c - c0] = (int)k.size();
    for (int s = 0; s < ns; ++s) {
      const TNode& nd = tree.nodes[L.ids[s]];
      const TNode& p = tree.nodes[nd.parent];
      const int sib_id = (p.left == L.ids[s]) ? p.right : p.left;
      L.sib[s] = tree.nodes[sib_id].pos;
      st.push_back(nd.start);
      ln.push_back(nd.len);
      sst.push_back(tree.nodes[sib_id].start);
      sln.push_back(tree.nodes[sib_id].len);
      lvl.push_back(c);
      pos.push_back(s);
      const int64_t m = nd.len, n = tree.nodes[sib_id].len;
      if (uses_rsvd(m, n, opt)) {
        mode.push_back(BlockMode::RSVD);
        k.push_back(sketch_width(m, n, opt));
      } else if (m <= n) {
        mode.push_back(BlockMode::ExactWide);
        k.push_back((int)m);
      } else {
        mode.push_back(BlockMode::ExactTall);
        k.push_back((int)n);
      }
    }
  }
  base[ncl] = (int)k.size();
  const int nt = (int)k.size();
  int kcap = 0;
  for (int kv : k) kcap = std::max(kcap, kv);
  HMAT_REQUIRE(kcap <= 300, HMAT_ERR_ARG, "block sketch width above 300 is not supported");
  const int64_t lvs = N * (int64_t)kcap;  // elements of one depth's column block
  auto cb = [&](int t) { return (int64_t)(lvl[t] - c0) * lvs; };
  const int64_t nk = lvs * ncl;
  Yb_.reserve(nk); Xb_.reserve(nk); Tb_.reserve(nk);
  zero_bytes_async(Yb_.get(), nk * sizeof(double), stream);
  zero_bytes_async(Xb_.get(), nk * sizeof(double), stream);
  const int64_t kk = (int64_t)kcap * kcap;
  Rb_.reserve(nt * kk); Vr_.reserve(nt * kk); Wk_.reserve(2 * nt * kk);
  TU_.reserve(nt * kk); TV_.reserve(nt * kk);
  Sv_.reserve((size_t)nt * kcap);
  info_.reserve(nt);

  // Omega (RSVD) / identity (exact) operands; a block's device Omega stream
  // is keyed by (depth, position), as when depths were built one at a time
  std::vector<SegNN> rnd, idx_col, idx_row;
  for (int t = 0; t < nt; ++t) {
    if (mode[t] == BlockMode::RSVD && !host_omega)
      rnd.push_back(SegNN{sst[t], (int64_t)lvl[t] * (1 << 24) + pos[t], sst[t] + cb(t), (int32_t)sln[t], 0, k[t], 0});
    if (mode[t] == BlockMode::ExactTall)
      idx_col.push_back(SegNN{0, 0, sst[t] + cb(t), (int32_t)sln[t], 0, k[t], 0});
    if (mode[t] == BlockMode::ExactWide)
      idx_row.push_back(SegNN{0, 0, st[t] + cb(t), (int32_t)ln[t], 0, k[t], 0});
  }
  HMAT_REQUIRE(!(host_omega && fixed_), HMAT_ERR_ARG, "a recorded build draws Omega on the device");
  if (host_omega) {
    bool any = false;
    for (int t = 0; t < nt; ++t) any |= (mode[t] == BlockMode::RSVD);
    if (any) {
      std::vector<double> h((size_t)nk, 0.0);
      for (int t = 0; t < nt; ++t) {
        if (mode[t] != BlockMode::RSVD) continue;
        const int64_t o = omega_off[lv[lvl[t]].ids[pos[t]]];
        HMAT_REQUIRE(o >= 0, HMAT_ERR_STATE, "omega schedule mismatch");
        for (int j = 0; j < k[t]; ++j)
          std::memcpy(&h[cb(t) + sst[t] + (int64_t)j * N], host_omega + o + (int64_t)j * sln[t],
                      sizeof(double) * sln[t]);
      }
      HIP_CHECK(hipMemcpyAsync(Xb_.get(), h.data(), nk * sizeof(double), hipMemcpyHostToDevice, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
    }
  } else {
    kern::fill_random_segs(rnd, dnn_, Xb_.get(), N, opt.seed, opt.omega_dist, stream);
  }
  kern::fill_identity_segs(idx_col, dnn_, Xb_.get(), N, stream);
  kern::fill_identity_segs(idx_row, dnn_, Yb_.get(), N, stream);

  // descriptors
  std::vector<SegSk> sk_fwd, sk_fwd_rsvd, sk_fwd_exact, sk_bwd_rsvd, sk_bwd;
  std::vector<OrthSeg> orow_rsvd, orow, ocol_rsvd, ocol;
  for (int t = 0; t < nt; ++t) {
    const int64_t o = cb(t);
    const SegSk fwd{st[t], sst[t], 0, sst[t] + o, st[t] + o, (int32_t)ln[t], (int32_t)sln[t], 0, k[t]};
    const SegSk bwd{sst[t], st[t], 0, st[t] + o, sst[t] + o, (int32_t)sln[t], (int32_t)ln[t], 0, k[t]};
    sk_bwd.push_back(bwd);
    ocol.push_back(OrthSeg{sst[t] + o, (int32_t)sln[t], k[t]});
    if (mode[t] != BlockMode::ExactWide) {
      sk_fwd.push_back(fwd);
      if (mode[t] != BlockMode::RSVD) sk_fwd_exact.push_back(fwd);
      orow.push_back(OrthSeg{st[t] + o, (int32_t)ln[t], k[t]});
    }
    if (mode[t] == BlockMode::RSVD) {
      sk_fwd_rsvd.push_back(fwd);
      sk_bwd_rsvd.push_back(bwd);
      orow_rsvd.push_back(OrthSeg{st[t] + o, (int32_t)ln[t], k[t]});
      ocol_rsvd.push_back(OrthSeg{sst[t] + o, (int32_t)sln[t], k[t]});
    }
  }

  // Y = A Omega, power iterations, Q = orth(Y)      (compression.h:78-84)
  // Mixed precision runs these range-finding passes on BF16 MFMA: the last
  // pass before Q with kLowPrecSplit terms, the power-iteration passes with
  // kPowerSplit.  The projection B^T = A^T Q below stays FP64 in both modes.
  const bool lp = (opt.precision == HMAT_PREC_MIXED);
  auto range_pass = [&](bool trans, const std::vector<SegSk>& sk, const double* X, double* Y, bool last) {
    if (lp)
      kern::sketch_lp(g, trans, last ? kLowPrecSplit : kPowerSplit, sk, X, N, Y, N, kcap, ws_, stream);
    else
      kern::sketch(g, trans, sk, X, N, Y, N, kcap, ws_, stream);
  };
  const int q = sk_fwd_rsvd.empty() ? 0 : opt.power_iters;
  if (lp && q > 0) {
    // the exact-branch blocks get no power iterations: their first pass is
    // their last
    range_pass(false, sk_fwd_rsvd, Xb_.get(), Yb_.get(), false);
    range_pass(false, sk_fwd_exact, Xb_.get(), Yb_.get(), true);
  } else {
    range_pass(false, sk_fwd, Xb_.get(), Yb_.get(), true);
  }
  // Intermediate bases of the mixed path come from single-BF16 passes, whose
  // rounding keeps them well conditioned (cond ~ 1e3): one shifted
  // CholeskyQR pass orthonormalises them to ~1e-10, far below their own error.
  auto orth_mid = [&](const std::vector<OrthSeg>& os, double* Y) {
    if (lp)
      orth_.run_once(os, Y, Tb_.get(), N, kcap, stream);
    else
      orth_.run(os, Y, Tb_.get(), N, kcap, nullptr, stream);
  };
  for (int it = 0; it < q; ++it) {
    orth_mid(orow_rsvd, Yb_.get());
    range_pass(true, sk_bwd_rsvd, Yb_.get(), Xb_.get(), false);
    orth_mid(ocol_rsvd, Xb_.get());
    range_pass(false, sk_fwd_rsvd, Xb_.get(), Yb_.get(), it == q - 1);
  }
  orth_.run(orow, Yb_.get(), Tb_.get(), N, kcap, nullptr, stream);
  // B^T = A^T Q (n x k), then B^T = Qb R                 (compression.h:87)
  kern::sketch(g, true, sk_bwd, Yb_.get(), N, Xb_.get(), N, kcap, ws_, stream);
  orth_.run(ocol, Xb_.get(), Tb_.get(), N, kcap, Rb_.get(), stream);
  // R = Ur S Vr^T  =>  A_blk ~ (Q Vr) S (Qb Ur)^T        (compression.h:90-98)
  std::vector<SmallMat> sm(nt);
  for (int t = 0; t < nt; ++t) sm[t] = SmallMat{(int64_t)t * kk, k[t], kcap};
  // the mixed-precision basis resolves the block to about 2^-23 of its
  // norm (FP32 entries in the last range pass): columns below that are not
  // resolved against each other
  kern::jacobi_svd(sm, dsm_, Rb_.get(), Vr_.get(), Sv_.get(), kcap, Wk_.get(), stream,
                   lp ? kMixedJacobiFloor : 1e-30);
  // a fixed-schedule build has fixed ranks and leaves its singular values
  // on the device (a non-finite one shows in the solve's residual)
  const std::vector<double> S =
      fixed_ ? std::vector<double>((size_t)nt * kcap, 0.0) : download(Sv_.get(), (size_t)nt * kcap, stream);

  // rank per block (compression.h:42-46 or fixed)
  std::vector<int> rank(nt, 0);
  int rmax_all = 0;
  for (int t = 0; t < nt; ++t) {
    const double* sv = &S[(size_t)t * kcap];
    int r = 0;
    if (opt.fixed_rank || fixed_) {
      r = std::min(opt.rank, k[t]);
    } else {
      if (k[t] > 0 && sv[0] > 1e-15)
        for (int i = 0; i < k[t]; ++i) r += (sv[i] > opt.tol * sv[0]) ? 1 : 0;
      if (r == 0) r = 1;
    }
    for (int i = 0; i < k[t]; ++i)
      if (!std::isfinite(sv[i])) throw Error(HMAT_ERR_NONFINITE, "non-finite singular value in build");
    rank[t] = r;
    rmax_all = std::max(rmax_all, r);
  }
  kern::trunc_prep(sm, dsm_, Rb_.get(), Vr_.get(), Sv_.get(), kcap, stage(rank, stream), rmax_all, TU_.get(),
                   TV_.get(), 0, stream);
  for (int c = c0; c <= c1; ++c) {
    HLevel& L = lv[c];
    const int b0 = base[c - c0], ns = base[c - c0 + 1] - b0;
    L.rank.assign(rank.begin() + b0, rank.begin() + b0 + ns);
    L.rmax = 0;
    for (int r : L.rank) L.rmax = std::max(L.rmax, r);
    L.S.assign((size_t)ns * L.rmax, 0.0);
    for (int s = 0; s < ns; ++s)
      for (int i = 0; i < L.rank[s]; ++i) L.S[(size_t)s * L.rmax + i] = S[(size_t)(b0 + s) * kcap + i];
    const int64_t nr = N * (int64_t)L.rmax;
    L.US.resize(nr);
    L.V.resize(nr);
    L.US.zero(stream);
    L.V.zero(stream);
    std::vector<SegNN> nu(ns), nv(ns);
    for (int s = 0; s < ns; ++s) {
      const int t = b0 + s;
      nu[s] = SegNN{st[t] + cb(t), (int64_t)t * kk, st[t], (int32_t)ln[t], k[t], L.rmax, 0};
      nv[s] = SegNN{sst[t] + cb(t), (int64_t)t * kk, sst[t], (int32_t)sln[t], k[t], L.rmax, 0};
    }
    kern::seg_gemm_nn(nu, Yb_.get(), N, TU_.get(), kcap, L.US.get(), N, 1.0, 0.0, stream);
    kern::seg_gemm_nn(nv, Xb_.get(), N, TV_.get(), kcap, L.V.get(), N, 1.0, 0.0, stream);
  }
}

// Y = D X over all leaves (leaf blocks stored lmax x lmax).
template <class F>
void HodlrEngine::leaf_apply_t(const F* Dm, const F* B, int64_t ldb, int ncols, F* Y, int64_t ldy, F alpha,
                               F beta) {
  const int64_t lm = tree.lmax;
  std::vector<SegNN> segs;
  for (size_t i = 0; i < tree.leaves.size(); ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    segs.push_back(SegNN{(int64_t)i * lm * lm, nd.start, nd.start, (int32_t)nd.len, (int32_t)nd.len, ncols, 0});
  }
  kern::seg_gemm_nn<F>(segs, Dm, lm, B, ldb, Y, ldy, alpha, beta, stream);
}

template <class F>
void HodlrEngine::leaf_solve_t(HFactor<F>& fa, const F* B, int64_t ldb, int ncols, F* Y, int64_t ldy, F* R) {
  leaf_apply_t<F>(fa.Dinv.get(), B, ldb, ncols, Y, ldy);
  if constexpr (kRefineBlocks<F>) {
    // R = B - D Y, Y += D^{-1} R
    if (R != B) kern::convert_cols<F, F>(N, ncols, B, ldb, R, N, stream);
    leaf_apply_t<F>(D.get(), Y, ldy, ncols, R, N, F(-1), F(1));
    leaf_apply_t<F>(fa.Dinv.get(), R, N, ncols, Y, ldy, F(1), F(1));
  }
}

void HodlrEngine::matvec(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy) {
  matvec_enqueue(dX, ldx, nrhs, dY, ldy);
  HIP_CHECK(hipStreamSynchronize(stream));
}

size_t HodlrEngine::matvec_scratch_need(int nrhs) const {
  const int nr = nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1);  // widest pass of matvec_with_scratch
  return ((size_t)mv_.nleaf * mv_.L * mv_.KM + (size_t)mv_.nseg * mv_.KM) * nr;
}

void HodlrEngine::matvec_with_scratch(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy,
                                      double* scratch, size_t cap) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "matvec before build");
  HMAT_REQUIRE(nrhs > 0 && ldx >= N && ldy >= N, HMAT_ERR_ARG, "bad matvec dimensions");
  HMAT_REQUIRE(kern::mv_lds_bytes(mv_, tree.lmax, 4) <= 64 * 1024, HMAT_ERR_ARG,
               "the operator's leaves are too large for the fused matvec kernels");
  const size_t need = matvec_scratch_need(nrhs);
  HMAT_REQUIRE(need <= cap, HMAT_ERR_ARG,
               "matvec needs " + std::to_string(need) + " doubles of scratch, the fixed scratch holds " +
                   std::to_string(cap));
  const int nr = nrhs >= 4 ? 4 : (nrhs >= 2 ? 2 : 1);
  double* part = scratch;
  double* t = scratch + (size_t)mv_.nleaf * mv_.L * mv_.KM * nr;
  const int64_t lm = tree.lmax;
  int done = 0;
  while (done < nrhs) {
    const int rem = nrhs - done;
    const int w = rem >= 4 ? 4 : (rem >= 2 ? 2 : 1);
    kern::hodlr_matvec_fused(mv_, D.get(), lm, dX + (int64_t)done * ldx, ldx, w, part, t,
                             dY + (int64_t)done * ldy, ldy, stream);
    done += w;
  }
  HIP_CHECK(hipStreamSynchronize(stream));
}

void HodlrEngine::matvec_enqueue(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "matvec before build");
  HMAT_REQUIRE(nrhs > 0 && ldx >= N && ldy >= N, HMAT_ERR_ARG, "bad matvec dimensions");
  // the fused passes for up to two passes of the widest width (16 columns,
  // 2^18 rows in leaves of 256: two 8-column passes 1.08 ms, the per-depth
  // products 1.80 ms); more columns share each read of the generators
  if (mv_width_ >= 4 && nrhs <= 2 * mv_width_) {
    ProfScope prof("matvec", stream, matvec_bytes(nrhs));
    matvec_async(dX, ldx, nrhs, dY, ldy);
    prof.finish();
    return;
  }
  ProfScope prof("matvec", stream, matvec_bytes(nrhs));
  leaf_apply_t<double>(D.get(), dX, ldx, nrhs, dY, ldy);
  for (int c = 1; c <= tree.max_depth; ++c) {
    HLevel& L = lv[c];
    const int ns = (int)L.ids.size();
    const int64_t tstride = (int64_t)L.rmax * nrhs;
    tvec_.reserve((size_t)ns * tstride);
    std::vector<SegTN> tn(ns);
    std::vector<SegNN> nn(ns);
    for (int s = 0; s < ns; ++s) {
      const TNode& nd = tree.nodes[L.ids[s]];
      // t_s = V_s^T x_s has rank(sib) entries
      tn[s] = SegTN{nd.start, nd.start, (int64_t)s * tstride, (int32_t)nd.len, L.rank[L.sib[s]], nrhs, 0};
      // y_s += (U S)_s t_sib
      nn[s] = SegNN{nd.start, (int64_t)L.sib[s] * tstride, nd.start, (int32_t)nd.len, L.rank[s], nrhs, 0};
    }
    kern::seg_gemm_tn<double>(tn, L.V.get(), N, dX, ldx, tvec_.get(), L.rmax, 1.0, 0.0, ws_, stream);
    kern::seg_gemm_nn<double>(nn, L.US.get(), N, tvec_.get(), L.rmax, dY, ldy, 1.0, 1.0, stream);
  }
  prof.finish();
}

// W(rows of internal depth-j nodes, 0:ncols) -= Q C M^{-1} V^T W, with
//   M = [[I, V_l^T Q_l], [V_r^T Q_r, I]] and C swapping the two halves.
template <class F>
void HodlrEngine::correction_t(HFactor<F>& fa, int j, F* W, int64_t ldw, int ncols, const kern::NodeTask* cap,
                               const kern::BGemm* capops) {
  HParents& P = par[j];
  const HLevel& C = lv[j + 1];
  const int np = (int)P.ids.size();
  if (np == 0 || ncols == 0) return;
  const int64_t mm = P.mmax;
  const int64_t tst = mm * ncols;
  fa.tvec.reserve((size_t)np * tst);
  fa.coef.reserve((size_t)np * tst);
  const F* Q = fa.Wall.get() + colofs[j + 1] * N;
  std::vector<SegTN> tn;
  std::vector<SegNN> mn, up;
  for (int p = 0; p < np; ++p) {
    const TNode& l = tree.nodes[C.ids[P.lpos[p]]];
    const TNode& r = tree.nodes[C.ids[P.rpos[p]]];
    const int rl = C.rank[P.lpos[p]], rr = C.rank[P.rpos[p]];
    tn.push_back(SegTN{l.start, l.start, (int64_t)p * tst, (int32_t)l.len, rr, ncols, 0});
    tn.push_back(SegTN{r.start, r.start, (int64_t)p * tst + rr, (int32_t)r.len, rl, ncols, 0});
    mn.push_back(SegNN{(int64_t)p * mm * mm, (int64_t)p * tst, (int64_t)p * tst, P.msize[p], P.msize[p], ncols, 0});
    up.push_back(SegNN{l.start, (int64_t)p * tst + rr, l.start, (int32_t)l.len, rl, ncols, 0});
    up.push_back(SegNN{r.start, (int64_t)p * tst, r.start, (int32_t)r.len, rr, ncols, 0});
  }
  kern::seg_gemm_tn<F>(tn, fa.V[j + 1], N, W, ldw, fa.tvec.get(), mm, F(1), F(0), ws_, stream);
  if (cap) {
    // the refined capacitance solves as node tasks (one workgroup per
    // parent, its three products in order): one launch instead of three
    kern::node_tasks(cap, np, capops, stream);
  } else {
    kern::seg_gemm_nn<F>(mn, fa.Minv[j].get(), mm, fa.tvec.get(), mm, fa.coef.get(), mm, F(1), F(0), stream);
    if constexpr (kRefineBlocks<F>) {
      // tvec <- tvec - M coef, coef += M^{-1} tvec
      kern::seg_gemm_nn<F>(mn, fa.Mmat[j].get(), mm, fa.coef.get(), mm, fa.tvec.get(), mm, F(-1), F(1), stream);
      kern::seg_gemm_nn<F>(mn, fa.Minv[j].get(), mm, fa.tvec.get(), mm, fa.coef.get(), mm, F(1), F(1), stream);
    }
  }
  kern::seg_gemm_nn<F>(up, Q, N, fa.coef.get(), mm, W, ldw, F(-1), F(1), stream);
}

template <class F>
static void copy_or_convert(const double* src, F* dst, int64_t n, hipStream_t s) {
  if constexpr (std::is_same<F, double>::value) {
    copy_bytes_async(dst, src, n * sizeof(double), s);
  } else {
    kern::convert<F, double>(src, dst, n, s);
  }
}

template <class F>
void HodlrEngine::factor_t(HFactor<F>& fa) {
  fa.clear();
  const int64_t lm = tree.lmax;
  const int nleaf = (int)tree.leaves.size();
  // leaf inverses (the reference refactorises D on every visit,
  // hodlr_routines_omp.h:150-154; here once)
  fa.Dinv.resize((size_t)nleaf * lm * lm);
  copy_or_convert<F>(D.get(), fa.Dinv.get(), (int64_t)nleaf * lm * lm, stream);
  std::vector<SmallMat> sm(nleaf);
  for (int i = 0; i < nleaf; ++i)
    sm[i] = SmallMat{(int64_t)i * lm * lm, (int32_t)tree.nodes[tree.leaves[i]].len, (int32_t)lm};
  // pivot status of every leaf and capacitance inversion, checked together
  // at the end (one synchronisation)
  int ninfo = nleaf;
  for (const TNode& nd : tree.nodes) ninfo += nd.leaf ? 0 : 1;
  info_.reserve(std::max(ninfo, 1));
  kern::gj_inverse<F>(sm, dsm_, fa.Dinv.get(), info_.get(), stream);
  std::vector<std::pair<int, int>> cap_info;  // (offset, count) per depth j, deepest first
  int info_off = nleaf;
  const int dmax = tree.max_depth;
  colofs.assign(dmax + 2, 0);
  for (int c = 1; c <= dmax; ++c) colofs[c + 1] = colofs[c] + lv[c].rmax;
  Rtot = colofs[dmax + 1];
  fa.V.assign(dmax + 1, nullptr);
  fa.Vown.resize(dmax + 1);
  fa.Minv.resize(std::max(dmax, 1));
  if constexpr (kRefineBlocks<F>) fa.Mmat.resize(std::max(dmax, 1));
  par.assign(std::max(dmax, 1), HParents{});
  auto check_info = [&]() {
    const auto info = download(info_.get(), info_off, stream);
    for (int i = 0; i < nleaf; ++i)
      if (info[i] != 0)
        throw Error(HMAT_ERR_SINGULAR, "singular leaf block " + std::to_string(i) + " (pivot " +
                                           std::to_string(info[i]) + ")");
    for (int jj = dmax - 1, q = 0; jj >= 0; --jj, ++q)
      for (int p = 0; p < cap_info[q].second; ++p)
        if (info[cap_info[q].first + p] != 0)
          throw Error(HMAT_ERR_SINGULAR, "singular capacitance matrix at depth " + std::to_string(jj));
  };
  last_ninfo_ = ninfo;
  if (dmax == 0) {
    if (!fixed_) check_info();
    fa.ready = true;
    return;
  }
  for (int c = 1; c <= dmax; ++c) {
    if constexpr (std::is_same<F, double>::value) {
      fa.V[c] = lv[c].V.get();
    } else {
      fa.Vown[c].resize(lv[c].V.size());
      kern::convert<F, double>(lv[c].V.get(), fa.Vown[c].get(), (int64_t)lv[c].V.size(), stream);
      fa.V[c] = fa.Vown[c].get();
    }
  }
  // W_c = D^{-1} (U S)_c for every depth c, stacked by columns
  fa.Wall.resize((size_t)N * Rtot);
  fa.wtmp.reserve((size_t)N * Rtot);
  for (int c = 1; c <= dmax; ++c)
    if (lv[c].rmax) copy_or_convert<F>(lv[c].US.get(), fa.wtmp.get() + colofs[c] * N, N * lv[c].rmax, stream);
  leaf_solve_t<F>(fa, fa.wtmp.get(), N, (int)Rtot, fa.Wall.get(), N, fa.wtmp.get());

  for (int j = dmax - 1; j >= 0; --j) {
    HParents& P = par[j];
    const HLevel& C = lv[j + 1];
    for (int id : tree.by_depth[j]) {
      const TNode& nd = tree.nodes[id];
      if (nd.leaf) continue;
      P.ids.push_back(id);
      P.lpos.push_back(tree.nodes[nd.left].pos);
      P.rpos.push_back(tree.nodes[nd.right].pos);
      P.msize.push_back(C.rank[tree.nodes[nd.left].pos] + C.rank[tree.nodes[nd.right].pos]);
    }
    const int np = (int)P.ids.size();
    P.mmax = 0;
    for (int m : P.msize) P.mmax = std::max(P.mmax, m);
    const int64_t mm = P.mmax;
    fa.Minv[j].resize((size_t)np * mm * mm);
    // M = [[I, V_l^T Q_l], [V_r^T Q_r, I]]
    std::vector<SegNN> eye(np);
    std::vector<SegTN> kb;
    std::vector<SmallMat> ms(np);
    const F* Q = fa.Wall.get() + colofs[j + 1] * N;
    for (int p = 0; p < np; ++p) {
      const TNode& l = tree.nodes[C.ids[P.lpos[p]]];
      const TNode& r = tree.nodes[C.ids[P.rpos[p]]];
      const int rl = C.rank[P.lpos[p]], rr = C.rank[P.rpos[p]];
      eye[p] = SegNN{0, 0, (int64_t)p * mm * mm, P.msize[p], 0, P.msize[p], 0};
      kb.push_back(SegTN{l.start, l.start, (int64_t)p * mm * mm + (int64_t)rr * mm, (int32_t)l.len, rr, rl, 0});
      kb.push_back(SegTN{r.start, r.start, (int64_t)p * mm * mm + rr, (int32_t)r.len, rl, rr, 0});
      ms[p] = SmallMat{(int64_t)p * mm * mm, P.msize[p], (int32_t)mm};
    }
    fa.Minv[j].zero(stream);
    kern::fill_identity_segs<F>(eye, dnn_, fa.Minv[j].get(), mm, stream);
    kern::seg_gemm_tn<F>(kb, fa.V[j + 1], N, Q, N, fa.Minv[j].get(), mm, F(1), F(0), ws_, stream);
    if constexpr (kRefineBlocks<F>) {
      fa.Mmat[j].resize((size_t)np * mm * mm);
      copy_bytes_async(fa.Mmat[j].get(), fa.Minv[j].get(), (size_t)np * mm * mm * sizeof(F), stream);
    }
    kern::gj_inverse<F>(ms, dsm_, fa.Minv[j].get(), info_.get() + info_off, stream);
    cap_info.push_back({info_off, np});
    info_off += np;
    // push the depth-j correction into the bases of shallower depths
    if (colofs[j + 1] > 0) correction_t<F>(fa, j, fa.Wall.get(), N, (int)colofs[j + 1]);
  }
  if (!fixed_) check_info();  // a fixed-schedule factorisation is checked by its caller
  fa.ready = true;
  plan_fused<F>(fa);
}

// Descriptors of the fused apply-inverse: per leaf its rows, per (leaf,
// parent depth) the parent and the column counts of V and Q at the node
// holding the leaf, per parent its leaf range and capacitance block.
template <class F>
void HodlrEngine::plan_fused(HFactor<F>& fa) {
  kern::ApPlan<F>& ap = fa.ap;
  ap.ok = false;
  const int L = tree.max_depth;
  const int nl = (int)tree.leaves.size();
  std::vector<kern::ApLeaf> lf(nl);
  std::vector<int64_t> starts(nl);
  for (int i = 0; i < nl; ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    lf[i] = kern::ApLeaf{nd.start, (int32_t)nd.len, 0};
    starts[i] = nd.start;
  }
  auto leaf_range = [&](const TNode& nd, int& l0, int& l1) {
    l0 = (int)(std::lower_bound(starts.begin(), starts.end(), nd.start) - starts.begin());
    l1 = (int)(std::lower_bound(starts.begin(), starts.end(), nd.start + nd.len) - starts.begin());
  };
  std::vector<kern::ApLevel<F>> levels(std::max(L, 1));
  std::vector<kern::ApEntry> ent((size_t)nl * std::max(L, 1), kern::ApEntry{-1, 0, 0, 0});
  std::vector<kern::ApParent> pars;
  int KM = 1, MM = 1;
  for (int j = 0; j < L; ++j) {
    const HParents& P = par[j];
    const HLevel& C = lv[j + 1];
    levels[j].V = fa.V[j + 1];
    levels[j].Q = fa.Wall.get() + colofs[j + 1] * N;
    levels[j].Minv = fa.Minv[j].get();
    if constexpr (kRefineBlocks<F>)
      levels[j].M = fa.Mmat[j].get();
    else
      levels[j].M = nullptr;
    levels[j].mm = P.mmax;
    KM = std::max(KM, C.rmax);
    MM = std::max(MM, P.mmax);
    for (size_t q = 0; q < P.ids.size(); ++q) {
      const int pg = (int)pars.size();
      const TNode& pn = tree.nodes[P.ids[q]];
      const TNode& l = tree.nodes[C.ids[P.lpos[q]]];
      const TNode& r = tree.nodes[C.ids[P.rpos[q]]];
      const int rl = C.rank[P.lpos[q]], rr = C.rank[P.rpos[q]];
      kern::ApParent ap_{(int64_t)q * P.mmax * P.mmax, 0, 0, 0, rr, rl, j, 0};
      int a0, a1, b0, b1;
      leaf_range(pn, a0, a1);
      leaf_range(l, b0, b1);
      ap_.leaf0 = a0;
      ap_.lsplit = b1;
      ap_.leaf1 = a1;
      pars.push_back(ap_);
      for (int i = b0; i < b1; ++i) ent[(size_t)i * L + j] = kern::ApEntry{pg, rr, rl, rr};
      leaf_range(r, b0, b1);
      for (int i = b0; i < b1; ++i) ent[(size_t)i * L + j] = kern::ApEntry{pg, rl, rr, 0};
    }
  }
  ap.nleaf = nl;
  ap.L = L;
  ap.KM = KM;
  ap.MM = MM;
  ap.npar = (int)pars.size();
  ap.lm = tree.lmax;
  ap.leaves.upload(lf, stream);
  ap.ent.upload(ent, stream);
  ap.lv.upload(levels, stream);
  if (!pars.empty()) ap.par.upload(pars, stream);
  ap.part.reserve((size_t)nl * KM * 4 * sizeof(F));
  ap.coef.reserve((size_t)std::max(ap.npar, 1) * MM * 4 * sizeof(F));
  ap.sync.reserve(((size_t)2 * ap.npar + 1) * sizeof(int));
  zero_bytes_async(ap.sync.template as<int>(), ((size_t)2 * ap.npar + 1) * sizeof(int), stream);
  ap.ok = true;
}

// The fused apply takes four columns per launch, and each workgroup walks
// its leaves one after another, each behind a barrier and an arrival
// counter.  It wins for a few columns over a few leaves per workgroup (one
// column, 1024 leaves of 256 rows: 0.90 ms against 2.17 ms); with many
// columns or many small leaves the wide launches of the multi-launch path
// win (64 columns over a forest of 20 480 leaves of 64 rows: 111 ms fused).
constexpr int kFusedMaxCols = 4;
constexpr int kFusedLeavesPerGroup = 4;

template <class F>
bool HodlrEngine::fused_ready(const HFactor<F>& fa, int ncols) const {
  if (fused_solve == 0 || !fa.ap.ok) return false;
  const int grid = kern::apply_fused_grid<F>(fa.ap, kRefineBlocks<F>, 4);
  if (grid <= 0) return false;
  if (fused_solve == 2) return true;
  return ncols <= kFusedMaxCols && (int64_t)fa.ap.nleaf <= (int64_t)kFusedLeavesPerGroup * grid;
}

void HodlrEngine::check_fused_active() {
  if (opt.precision == HMAT_PREC_MIXED)
    check_fused<float>(f32_);
  else
    check_fused<double>(f64_);
}

template <class F>
void HodlrEngine::check_fused(HFactor<F>& fa) {
  if (!fa.ap.ok) return;
  int err = 0;
  HIP_CHECK(hipMemcpyAsync(&err, fa.ap.sync.template as<int>() + 2 * fa.ap.npar, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  if (err) {
    zero_bytes_async(fa.ap.sync.template as<int>() + 2 * fa.ap.npar, sizeof(int), stream);
    throw Error(HMAT_ERR_DEVICE, "fused solve: a wait of the persistent kernel exceeded its spin limit");
  }
}

// The refined capacitance solves of an FP64 apply, coef = M^{-1} t,
// t -= M coef, coef += M^{-1} t, as node tasks (one workgroup per parent
// running its three products in order) at the parent depths where that is
// the cheaper form: one launch per depth instead of three segmented
// products, each of which is a ~15 us launch at rank 32 and 16 columns
// whatever the number of parents.  A workgroup's products run on the vector
// units, so with many parents or wide column blocks the segmented MFMA
// products win (64 columns over 7481 parents of rank 24: 1.34 ms per depth
// as node tasks); those depths get first[j] = -1.  tvec and coef must
// already hold np x mmax x ncols values at every depth.
constexpr int kCapTaskCols = 16, kCapTaskParents = 512;
static void cap_program(const std::vector<HParents>& par, const HFactor<double>& fa, int ncols,
                        std::vector<kern::BGemm>& ops, std::vector<kern::NodeTask>& tasks, std::vector<int>& first) {
  ops.clear();
  tasks.clear();
  first.assign(std::max<size_t>(par.size(), 1), -1);
  for (int j = (int)par.size() - 1; j >= 0; --j) {
    const HParents& P = par[j];
    if (ncols > kCapTaskCols || P.ids.size() > (size_t)kCapTaskParents) continue;
    const int mm = P.mmax;
    const int64_t tst = (int64_t)mm * ncols;
    first[j] = (int)tasks.size();
    for (size_t p = 0; p < P.ids.size(); ++p) {
      const double* Mi = fa.Minv[j].get() + (int64_t)p * mm * mm;
      const double* Mm = fa.Mmat[j].get() + (int64_t)p * mm * mm;
      double* t = fa.tvec.get() + (int64_t)p * tst;
      double* cf = fa.coef.get() + (int64_t)p * tst;
      const int ms = P.msize[p];
      tasks.push_back(kern::NodeTask{(int32_t)ops.size(), 3});
      ops.push_back(kern::BGemm{Mi, t, cf, ms, ncols, ms, mm, mm, mm, 0, 0, 1.0, 0.0});
      ops.push_back(kern::BGemm{Mm, cf, t, ms, ncols, ms, mm, mm, mm, 0, 0, -1.0, 1.0});
      ops.push_back(kern::BGemm{Mi, t, cf, ms, ncols, ms, mm, mm, mm, 0, 0, 1.0, 1.0});
    }
  }
}

template <class F>
void HodlrEngine::apply_inverse_t(HFactor<F>& fa, const F* B, int64_t ldb, int ncols, F* X,
                                  int64_t ldx) {
  if (fused_ready(fa, ncols)) {
    const F* Dref = nullptr;
    if constexpr (kRefineBlocks<F>) Dref = D.get();
    if (kern::apply_inverse_fused<F>(fa.ap, fa.Dinv.get(), Dref, B, ldb, ncols, X, ldx, N, stream)) return;
    fa.ap.ok = false;  // refused by the runtime: the multi-launch path from now on
  }
  if constexpr (kRefineBlocks<F>) fa.rtmp.reserve((size_t)N * ncols);
  leaf_solve_t<F>(fa, B, ldb, ncols, X, ldx, fa.rtmp.get());
  if constexpr (kRefineBlocks<F>) {
    // the capacitance solves as node tasks where cap_program takes that
    // form, all depths' descriptors staged in one copy (the recorded solve
    // graph runs the same tasks)
    const int dmax = tree.max_depth;
    size_t need = 1;
    for (int j = 0; j < dmax; ++j) need = std::max(need, par[j].ids.size() * (size_t)par[j].mmax * ncols);
    fa.tvec.reserve(need);
    fa.coef.reserve(need);
    std::vector<kern::BGemm> ops;
    std::vector<kern::NodeTask> tasks;
    std::vector<int> first;
    cap_program(par, fa, ncols, ops, tasks, first);
    StageGroup g;
    const int io = g.add(ops), it = g.add(tasks);
    g.put(stream);
    for (int j = dmax - 1; j >= 0; --j)
      correction_t<F>(fa, j, X, ldx, ncols, first[j] >= 0 ? g.get<kern::NodeTask>(it) + first[j] : nullptr,
                      g.get<kern::BGemm>(io));
    return;
  }
  for (int j = tree.max_depth - 1; j >= 0; --j) correction_t<F>(fa, j, X, ldx, ncols);
}

void HodlrEngine::factor() {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "factor before build");
  ProfScope prof("factor", stream, 0.0);
  if (opt.precision == HMAT_PREC_MIXED) {
    f64_.clear();
    factor_t<float>(f32_);
  } else {
    f32_.clear();
    factor_t<double>(f64_);
  }
  prof.finish();
  factored = true;
}

// One application of the factored inverse, FP64 in and out.
// Whether this apply-inverse replays the recorded graph (recording it first
// when needed): only once kEagerSolves eager applies of the column count ran.
template <class F>
static bool graph_turn(SolvePlan<F>& sp, bool enabled, int ncols) {
  if (!enabled) return false;
  if (sp.ncols == ncols && sp.graph.e) return true;
  return ++sp.uses[ncols] > kEagerSolves;
}

// Bytes one apply-inverse must move at least: the leaf inverses (and, for
// the refined FP64 leaf solve, the leaf blocks), Q and V of every depth, the
// right-hand side in and the solution out in FP64.
double HodlrEngine::solve_bytes(int ncols) const {
  const bool mixed = opt.precision == HMAT_PREC_MIXED;
  const double fb = mixed ? 4.0 : 8.0;
  double leaf2 = 0.0;
  for (int id : tree.leaves) leaf2 += (double)tree.nodes[id].len * (double)tree.nodes[id].len;
  return fb * ((mixed ? 1.0 : 2.0) * leaf2 + 2.0 * (double)N * (double)Rtot) + 16.0 * (double)N * ncols;
}

void HodlrEngine::apply_inverse(const double* B, int64_t ldb, int ncols, double* X, int64_t ldx) {
  ProfScope prof("solve_apply", stream, solve_bytes(ncols));
  apply_inverse_impl(B, ldb, ncols, X, ldx);
  prof.finish();
}

void HodlrEngine::apply_inverse_impl(const double* B, int64_t ldb, int ncols, double* X, int64_t ldx) {
  if (opt.precision != HMAT_PREC_MIXED) {
    // the fused apply is one launch per four columns: no graph to record
    if (!graph_turn(f64_.plan, solve_graph && !fused_ready(f64_, ncols), ncols)) {
      apply_inverse_t<double>(f64_, B, ldb, ncols, X, ldx);
      return;
    }
    plan_solve<double>(f64_, ncols);
    kern::convert_cols<double, double>(N, ncols, B, ldb, f64_.b.get(), N, stream);
    run_solve_graph<double>(f64_, ncols);
    kern::convert_cols<double, double>(N, ncols, f64_.x.get(), N, X, ldx, stream);
    return;
  }
  const bool graph = graph_turn(f32_.plan, solve_graph && !fused_ready(f32_, ncols), ncols);
  if (graph) plan_solve<float>(f32_, ncols);
  f32_.b.reserve((size_t)N * ncols);
  f32_.x.reserve((size_t)N * ncols);
  kern::convert_cols<float, double>(N, ncols, B, ldb, f32_.b.get(), N, stream);
  if (graph)
    run_solve_graph<float>(f32_, ncols);
  else
    apply_inverse_t<float>(f32_, f32_.b.get(), N, ncols, f32_.x.get(), N);
  kern::convert_cols<double, float>(N, ncols, f32_.x.get(), N, X, ldx, stream);
}

// Every device buffer that the recorded apply-inverse reads or writes.
template <class F>
static std::vector<const void*> solve_buffers(const HFactor<F>& fa) {
  std::vector<const void*> v{fa.Dinv.get(), fa.Wall.get(), fa.tvec.get(), fa.coef.get(),
                             fa.b.get(), fa.x.get(), fa.plan.part.get(), fa.rtmp.get()};
  for (const auto& m : fa.Minv) v.push_back(m.get());
  for (const auto& m : fa.Mmat) v.push_back(m.get());
  for (const F* p : fa.V) v.push_back(p);
  return v;
}

// Plan every launch of the apply-inverse for `ncols` columns and record them
// into a graph (the same launches, in the same order, as apply_inverse_t).
// A graph recorded earlier is kept only while every buffer it uses is still
// at its recorded address.
template <class F>
void HodlrEngine::plan_solve(HFactor<F>& fa, int ncols) {
  SolvePlan<F>& sp = fa.plan;
  if (sp.ncols == ncols && sp.graph.e && sp.bufs == solve_buffers(fa)) return;
  sp.clear();
  const int64_t lm = tree.lmax;
  fa.b.reserve((size_t)N * ncols);
  fa.x.reserve((size_t)N * ncols);
  {
    std::vector<SegNN> segs;
    for (size_t i = 0; i < tree.leaves.size(); ++i) {
      const TNode& nd = tree.nodes[tree.leaves[i]];
      segs.push_back(SegNN{(int64_t)i * lm * lm, nd.start, nd.start, (int32_t)nd.len, (int32_t)nd.len, ncols, 0});
    }
    kern::plan_nn(segs, sp.leaf, stream);
  }
  const int dmax = tree.max_depth;
  sp.tn.resize(std::max(dmax, 1));
  sp.mn.resize(std::max(dmax, 1));
  sp.up.resize(std::max(dmax, 1));
  sp.mm.assign(std::max(dmax, 1), 0);
  size_t tv = 1, slots = 1;
  for (int j = dmax - 1; j >= 0; --j) {
    HParents& P = par[j];
    const HLevel& C = lv[j + 1];
    const int np = (int)P.ids.size();
    if (np == 0) continue;
    const int64_t mm = P.mmax;
    const int64_t tst = mm * ncols;
    sp.mm[j] = mm;
    tv = std::max(tv, (size_t)np * tst);
    std::vector<SegTN> tn;
    std::vector<SegNN> mn, up;
    for (int p = 0; p < np; ++p) {
      const TNode& l = tree.nodes[C.ids[P.lpos[p]]];
      const TNode& r = tree.nodes[C.ids[P.rpos[p]]];
      const int rl = C.rank[P.lpos[p]], rr = C.rank[P.rpos[p]];
      tn.push_back(SegTN{l.start, l.start, (int64_t)p * tst, (int32_t)l.len, rr, ncols, 0});
      tn.push_back(SegTN{r.start, r.start, (int64_t)p * tst + rr, (int32_t)r.len, rl, ncols, 0});
      mn.push_back(SegNN{(int64_t)p * mm * mm, (int64_t)p * tst, (int64_t)p * tst, P.msize[p], P.msize[p], ncols, 0});
      up.push_back(SegNN{l.start, (int64_t)p * tst + rr, l.start, (int32_t)l.len, rl, ncols, 0});
      up.push_back(SegNN{r.start, (int64_t)p * tst, r.start, (int32_t)r.len, rr, ncols, 0});
    }
    kern::plan_tn(tn, sp.tn[j], stream);
    kern::plan_nn(mn, sp.mn[j], stream);
    kern::plan_nn(up, sp.up[j], stream);
    slots = std::max(slots, (size_t)sp.tn[j].nslots);
  }
  fa.tvec.reserve(tv);
  fa.coef.reserve(tv);
  if constexpr (kRefineBlocks<F>) {
    fa.rtmp.reserve((size_t)N * ncols);
    std::vector<kern::BGemm> ops;
    std::vector<kern::NodeTask> tasks;
    cap_program(par, fa, ncols, ops, tasks, sp.cap_first);
    if (!ops.empty()) {
      sp.cap_ops.upload(ops, stream);
      sp.cap_tasks.upload(tasks, stream);
    }
  }
  sp.part.reserve(slots * 4096);
  HIP_CHECK(hipStreamSynchronize(stream));
  // record
  HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  kern::launch_nn<F>(sp.leaf, fa.Dinv.get(), lm, fa.b.get(), N, fa.x.get(), N, F(1), F(0), stream);
  if constexpr (kRefineBlocks<F>) {
    kern::convert_cols<F, F>(N, ncols, fa.b.get(), N, fa.rtmp.get(), N, stream);
    kern::launch_nn<F>(sp.leaf, D.get(), lm, fa.x.get(), N, fa.rtmp.get(), N, F(-1), F(1), stream);
    kern::launch_nn<F>(sp.leaf, fa.Dinv.get(), lm, fa.rtmp.get(), N, fa.x.get(), N, F(1), F(1), stream);
  }
  for (int j = dmax - 1; j >= 0; --j) {
    if (par[j].ids.empty()) continue;
    const int64_t mm = sp.mm[j];
    const F* Q = fa.Wall.get() + colofs[j + 1] * N;
    kern::launch_tn<F>(sp.tn[j], fa.V[j + 1], N, fa.x.get(), N, fa.tvec.get(), mm, F(1), F(0), sp.part.get(),
                       stream);
    bool tasks = false;
    if constexpr (kRefineBlocks<F>) tasks = sp.cap_first[j] >= 0;
    if (tasks) {
      kern::node_tasks(sp.cap_tasks.get() + sp.cap_first[j], (int)par[j].ids.size(), sp.cap_ops.get(), stream);
    } else {
      kern::launch_nn<F>(sp.mn[j], fa.Minv[j].get(), mm, fa.tvec.get(), mm, fa.coef.get(), mm, F(1), F(0), stream);
      if constexpr (kRefineBlocks<F>) {
        kern::launch_nn<F>(sp.mn[j], fa.Mmat[j].get(), mm, fa.coef.get(), mm, fa.tvec.get(), mm, F(-1), F(1),
                           stream);
        kern::launch_nn<F>(sp.mn[j], fa.Minv[j].get(), mm, fa.tvec.get(), mm, fa.coef.get(), mm, F(1), F(1),
                           stream);
      }
    }
    kern::launch_nn<F>(sp.up[j], Q, N, fa.coef.get(), mm, fa.x.get(), N, F(-1), F(1), stream);
  }
  HIP_CHECK(hipStreamEndCapture(stream, &sp.graph.g));
  HIP_CHECK(hipGraphInstantiate(&sp.graph.e, sp.graph.g, nullptr, nullptr, 0));
  sp.ncols = ncols;
  sp.bufs = solve_buffers(fa);
}

template <class F>
void HodlrEngine::run_solve_graph(HFactor<F>& fa, int ncols) {
  HMAT_REQUIRE(fa.plan.ncols == ncols && fa.plan.graph.e, HMAT_ERR_STATE, "solve graph not planned");
  HIP_CHECK(hipGraphLaunch(fa.plan.graph.e, stream));
}
template void HodlrEngine::plan_solve<double>(HFactor<double>&, int);
template void HodlrEngine::plan_solve<float>(HFactor<float>&, int);
template void HodlrEngine::run_solve_graph<double>(HFactor<double>&, int);
template void HodlrEngine::run_solve_graph<float>(HFactor<float>&, int);

// X = A_H^{-1} B; then FP64 iterative refinement on the compressed operator
// (residual through the FP64 matvec, correction accumulated in FP64) until
// max_j ||b_j - A_H x_j|| / ||b_j|| <= refine_rtol.
void HodlrEngine::solve(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "solve before build");
  HMAT_REQUIRE(nrhs > 0 && ldb >= N && ldx >= N, HMAT_ERR_ARG, "bad solve dimensions");
  if (!factored) factor();
  apply_inverse(dB, ldb, nrhs, dX, ldx);
  refine_loop(dB, ldb, nrhs, dX, ldx);
}

void HodlrEngine::refine_loop(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx) {
  last_iters = 0;
  last_resid = -1.0;
  if (!(refine_rtol > 0.0)) {
    HIP_CHECK(hipStreamSynchronize(stream));
    check_fused_active();
    return;
  }
  ir_y_.reserve((size_t)N * nrhs);
  ir_r_.reserve((size_t)N * nrhs);
  ir_n_.reserve((size_t)2 * nrhs);
  ir_p_.reserve((size_t)64 * nrhs);
  kern::col_norms2(dB, ldb, N, nrhs, ir_n_.get(), ir_p_.get(), stream);
  auto bn = download(ir_n_.get(), nrhs, stream);
  double resid = 0.0, prev = 0.0;
  int it = 0;
  const bool mixed = opt.precision == HMAT_PREC_MIXED;
  ir_best_.reserve((size_t)N * nrhs);
  for (;; ++it) {
    matvec_enqueue(dX, ldx, nrhs, ir_y_.get(), N);
    kern::axpby_cols(N, nrhs, 1.0, dB, ldb, -1.0, ir_y_.get(), N, ir_r_.get(), N, stream);
    kern::col_norms2(ir_r_.get(), N, N, nrhs, ir_n_.get(), ir_p_.get(), stream);
    auto rn = download(ir_n_.get(), nrhs, stream);
    resid = 0.0;
    for (int j = 0; j < nrhs; ++j) {
      const double rj = bn[j] > 0.0 ? std::sqrt(rn[j] / bn[j]) : std::sqrt(rn[j]);
      // (std::max would drop a NaN.)  In mixed precision a non-finite
      // residual means the FP32 factorisation cannot solve this system.
      if (!std::isfinite(rj))
        throw Error(mixed ? HMAT_ERR_NOT_CONVERGED : HMAT_ERR_NONFINITE,
                    "non-finite residual after " + std::to_string(it) + " refinement iterations" +
                        (mixed ? " (the FP32 factorisation is too inaccurate for this system)" : ""));
      resid = std::max(resid, rj);
    }
    // a correction that did not reduce the residual: the iteration diverges;
    // X goes back to the previous iterate, the one with the lowest residual
    if (it > 0 && !(resid < prev)) {
      HIP_CHECK(hipMemcpy2DAsync(dX, ldx * sizeof(double), ir_best_.get(), N * sizeof(double),
                                 N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream));
      resid = prev;
      --it;
      break;
    }
    if (resid <= refine_rtol || it >= refine_max) break;
    prev = resid;
    // the iterate before the correction (every earlier one had a larger residual)
    HIP_CHECK(hipMemcpy2DAsync(ir_best_.get(), N * sizeof(double), dX, ldx * sizeof(double),
                               N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream));
    // correction d = A_H^{-1} r, accumulated into X in FP64
    apply_inverse(ir_r_.get(), N, nrhs, ir_y_.get(), N);
    kern::axpby_cols(N, nrhs, 1.0, dX, ldx, 1.0, ir_y_.get(), N, dX, ldx, stream);
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  check_fused_active();
  last_iters = it;
  last_resid = resid;
  // the target was asked for (by default in mixed precision, opt-in in FP64)
  // and missed: X holds the iterate with the lowest residual, and the call
  // says so
  if (resid > refine_rtol)
    throw Error(HMAT_ERR_NOT_CONVERGED, "refinement stopped at relative residual " +
                                            sci(resid) + " after " + std::to_string(it) +
                                            " iterations (target " + sci(refine_rtol) + ")");
}

// ---------------------------------------------------------------------------
// Batches recorded into one graph

void HodlrEngine::fixed_solve(int ncols, int& iters, bool adaptive, double& resid) {
  double* B = bg_.b.get();
  double* X = bg_.x_.get();
  apply_inverse(B, N, ncols, X, N);
  ir_y_.reserve((size_t)N * ncols);
  ir_r_.reserve((size_t)N * ncols);
  ir_n_.reserve((size_t)2 * ncols);
  ir_p_.reserve((size_t)64 * ncols);
  kern::col_norms2(B, N, N, ncols, ir_n_.get(), ir_p_.get(), stream);
  std::vector<double> bn;
  if (adaptive) bn = download(ir_n_.get(), ncols, stream);
  resid = 0.0;
  double prev = 0.0;
  int it = 0;
  for (;; ++it) {
    matvec_enqueue(X, N, ncols, ir_y_.get(), N);
    kern::axpby_cols(N, ncols, 1.0, B, N, -1.0, ir_y_.get(), N, ir_r_.get(), N, stream);
    if (adaptive) {
      // the refinement count that this batch takes (recorded with the graph)
      kern::col_norms2(ir_r_.get(), N, N, ncols, ir_n_.get() + ncols, ir_p_.get(), stream);
      const auto rn = download(ir_n_.get() + ncols, ncols, stream);
      resid = 0.0;
      for (int j = 0; j < ncols; ++j) {
        const double rj = bn[j] > 0.0 ? std::sqrt(rn[j] / bn[j]) : std::sqrt(rn[j]);
        resid = std::isfinite(rj) ? std::max(resid, rj) : HUGE_VAL;
      }
      if (!(refine_rtol > 0.0) || resid <= refine_rtol || it >= refine_max || (it > 0 && !(resid < prev))) break;
      prev = resid;
    } else if (it >= iters) {
      kern::col_norms2(ir_r_.get(), N, N, ncols, ir_n_.get() + ncols, ir_p_.get(), stream);
      break;
    }
    apply_inverse(ir_r_.get(), N, ncols, ir_y_.get(), N);
    kern::axpby_cols(N, ncols, 1.0, X, N, 1.0, ir_y_.get(), N, X, N, stream);
  }
  iters = it;
}

std::string HodlrEngine::batch_check(int ncols, double& resid) {
  std::string why = orth_.check_fixed(bg_.orth_calls, stream);
  if (!why.empty()) return why;
  const auto info = download(info_.get(), (size_t)bg_.ninfo, stream);
  for (int v : info)
    if (v != 0) return "singular leaf block or capacitance matrix";
  const auto n = download(ir_n_.get(), (size_t)2 * ncols, stream);
  resid = 0.0;
  for (int j = 0; j < ncols; ++j) {
    const double rj = n[j] > 0.0 ? std::sqrt(n[ncols + j] / n[j]) : std::sqrt(n[ncols + j]);
    if (!std::isfinite(rj)) return "non-finite residual";
    resid = std::max(resid, rj);
  }
  return std::string();
}

std::vector<const void*> HodlrEngine::batch_buffers() const {
  std::vector<const void*> v{Yb_.get(), Xb_.get(), Tb_.get(), Rb_.get(), Vr_.get(), Sv_.get(), Wk_.get(),
                             TU_.get(), TV_.get(), tvec_.get(), info_.get(), D.get(), dnn_.get(), dsm_.get(),
                             dbd_.get(), mv_leaves_.get(), mv_lvls_.get(), mv_ent_.get(), mv_segs_.get(),
                             mv_part_.get(), mv_t_.get(), ir_y_.get(), ir_r_.get(), ir_n_.get(), ir_p_.get(),
                             f64_.Dinv.get(), f64_.Wall.get(), f64_.wtmp.get(), f64_.tvec.get(), f64_.coef.get(),
                             f64_.rtmp.get(), f64_.b.get(), f64_.x.get(), bg_.b.get(), bg_.x_.get(),
                             bg_.shift.get()};
  for (const auto& m : f64_.Minv) v.push_back(m.get());
  for (const auto& m : f64_.Mmat) v.push_back(m.get());
  for (size_t c = 1; c < lv.size(); ++c) {
    v.push_back(lv[c].US.get());
    v.push_back(lv[c].V.get());
  }
  f64_.ap.addresses(v);
  ws_.addresses(v);
  orth_.addresses(v);
  return v;
}

namespace {
// the kernels whose parameter p1 is a diagonal shift (p1 on i == j only)
bool has_diag_shift(GenKind k) {
  return k == GenKind::Gaussian || k == GenKind::RefCauchy || k == GenKind::RefGaussian ||
         k == GenKind::RefCoulomb;
}
}  // namespace

int HodlrEngine::batch_solve(const GenDesc& g, const std::vector<int64_t>& sizes, int leaf_size, const double* dB,
                             int64_t ldb, int ncols, double* dX, int64_t ldx, bool use_graph) {
  int64_t ntot = 0;
  for (int64_t n : sizes) ntot += n;
  HMAT_REQUIRE(ncols > 0 && ldb >= ntot && ldx >= ntot, HMAT_ERR_ARG, "bad batch solve dimensions");
  if (!use_graph) {
    build_forest(g, sizes, leaf_size);
    factor();
    solve(dB, ldb, ncols, dX, ldx);
    return kBatchEager;
  }
  HMAT_REQUIRE(opt.fixed_rank && opt.precision == HMAT_PREC_FP64, HMAT_ERR_ARG,
               "a recorded batch needs fixed ranks and FP64");
  HMAT_REQUIRE(g.kind != GenKind::Dense, HMAT_ERR_ARG, "a recorded batch generates its entries");
  const double shift = has_diag_shift(g.kind) ? g.p1 : 0.0;
  GenDesc g0 = g;
  if (has_diag_shift(g.kind)) g0.p1 = 0.0;
  auto same_opts = [](const hmat_options& a, const hmat_options& b) {
    return a.rank == b.rank && a.fixed_rank == b.fixed_rank && a.oversample == b.oversample &&
           a.power_iters == b.power_iters && a.exact_below == b.exact_below && a.precision == b.precision &&
           a.omega_dist == b.omega_dist && a.seed == b.seed;
  };
  batch_note.clear();
  bool redo = false;
  const bool same = bg_.valid && bg_.graph.e && bg_.sizes == sizes && bg_.leaf == leaf_size &&
                    bg_.ncols == ncols && bg_.kind == (int)g.kind && bg_.x == g.x && bg_.y == g.y &&
                    bg_.p0 == g.p0 && same_opts(bg_.opt, opt) && bg_.bufs == batch_buffers();
  if (same) {
    kern::set_scalar(bg_.shift.get(), shift, stream);
    kern::convert_cols<double, double>(N, ncols, dB, ldb, bg_.b.get(), N, stream);
    HIP_CHECK(hipGraphLaunch(bg_.graph.e, stream));
    kern::convert_cols<double, double>(N, ncols, bg_.x_.get(), N, dX, ldx, stream);
    double resid = 0.0;
    const std::string why = batch_check(ncols, resid);
    if (why.empty()) {
      last_iters = bg_.iters;
      last_resid = resid;
      if (!(refine_rtol > 0.0) || resid <= refine_rtol) return kBatchReplayed;
      // the recorded refinement count fell short for these values: refine on
      refine_loop(dB, ldb, ncols, dX, ldx);
      return kBatchRedone;
    }
    // a status failed on these values (typically a basis that needs another
    // CholeskyQR pass): the eager pass below redoes the work, choosing the
    // pass counts for these values, and records them
    batch_note = "replay rejected: " + why;
    redo = true;
  }

  // First call for this shape: the fixed schedule eagerly, recording its
  // descriptors; then the same schedule recorded into the graph.
  bg_.drop();
  bg_.sizes = sizes;
  bg_.leaf = leaf_size;
  bg_.ncols = ncols;
  bg_.kind = (int)g.kind;
  bg_.x = g.x;
  bg_.y = g.y;
  bg_.p0 = g.p0;
  bg_.opt = opt;
  bg_.b.reserve((size_t)ntot * ncols);
  bg_.x_.reserve((size_t)ntot * ncols);
  bg_.shift.reserve(1);
  bg_.log.reset(new StageLog());
  kern::set_scalar(bg_.shift.get(), shift, stream);
  kern::convert_cols<double, double>(ntot, ncols, dB, ldb, bg_.b.get(), ntot, stream);
  // the fixed-schedule state for the duration of one pass (restored on any exit)
  struct Fixed {
    HodlrEngine& e;
    int fused;
    bool graph;
    bool capture;
    Fixed(HodlrEngine& en, bool cap) : e(en), fused(en.fused_solve), graph(en.solve_graph), capture(cap) {
      e.fixed_ = true;
      e.capturing_ = cap;
      e.diag_shift_ = e.bg_.shift.get();
      e.orth_.set_fixed(true, !cap);  // the eager pass chooses the pass counts, the recording repeats them
      e.fused_solve = 0;       // multi-launch applies (recordable, any column count)
      e.solve_graph = false;   // no nested solve graph
      stager(e.stream).attach(e.bg_.log.get());
      autotune_suspend(true);
      if (cap) prof_suspend(true);
    }
    ~Fixed() {
      e.fixed_ = false;
      e.capturing_ = false;
      e.diag_shift_ = nullptr;
      e.orth_.set_fixed(false);
      e.fused_solve = fused;
      e.solve_graph = graph;
      stager(e.stream).attach(nullptr);
      autotune_suspend(false);
      if (capture) prof_suspend(false);
    }
  };
  int iters = 0;
  double resid = 0.0;
  {
    bg_.log->mode = StageLog::Persist;
    Fixed fx(*this, false);
    build_forest(g0, sizes, leaf_size);
    factor();
    fixed_solve(ncols, iters, true, resid);
    bg_.orth_calls = orth_.fixed_calls();
    bg_.ninfo = last_ninfo_;
    bg_.iters = iters;
  }
  kern::convert_cols<double, double>(N, ncols, bg_.x_.get(), N, dX, ldx, stream);
  std::string why = batch_check(ncols, resid);
  if (!why.empty()) {
    // this batch needs the adaptive schedule: no graph
    batch_note = "fixed schedule failed: " + why;
    bg_.drop();
    build_forest(g, sizes, leaf_size);
    factor();
    solve(dB, ldb, ncols, dX, ldx);
    return redo ? kBatchRedone : kBatchEager;
  }
  last_iters = iters;
  last_resid = resid;
  if (refine_rtol > 0.0 && resid > refine_rtol) {
    // the fixed schedule does not reach the target here: no graph
    batch_note = "the fixed schedule stops at residual " + sci(resid);
    refine_loop(dB, ldb, ncols, dX, ldx);
    bg_.drop();
    return redo ? kBatchRedone : kBatchEager;
  }
  // record the same launches (nothing runs; the device keeps this call's results)
  bool recorded = false;
  {
    bg_.log->mode = StageLog::Replay;
    bg_.log->next = 0;
    Fixed fx(*this, true);
    HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      build_forest(g0, sizes, leaf_size);
      factor();
      int it = bg_.iters;
      double r = 0.0;
      fixed_solve(ncols, it, false, r);
      HIP_CHECK(hipStreamEndCapture(stream, &bg_.graph.g));
      recorded = !bg_.log->failed && bg_.log->next == bg_.log->dev.size() &&
                 orth_.fixed_calls() == bg_.orth_calls && last_ninfo_ == bg_.ninfo;
      if (!recorded)
        batch_note = "recording differs from the eager pass (descriptors " + std::to_string(bg_.log->next) + "/" +
                     std::to_string(bg_.log->dev.size()) + (bg_.log->failed ? ", mismatch" : "") + ")";
    } catch (const std::exception& ex) {  // the stream leaves capture mode; this call's result stands
      batch_note = std::string("recording failed: ") + ex.what();
      hipGraph_t gg = nullptr;
      (void)hipStreamEndCapture(stream, &gg);
      if (gg) (void)hipGraphDestroy(gg);
      (void)hipGetLastError();
      recorded = false;
    }
  }
  if (recorded) {
    HIP_CHECK(hipGraphInstantiate(&bg_.graph.e, bg_.graph.g, nullptr, nullptr, 0));
    bg_.bufs = batch_buffers();
  }
  bg_.valid = recorded;
  if (!recorded) {
    bg_.drop();
    prof_add("batch_graph_refused", 0.0, 1.0);
  }
  built = factored = true;
  if (redo) return kBatchRedone;
  return recorded ? kBatchRecorded : kBatchEager;
}

void HodlrEngine::to_dense_host(double* A, int64_t lda) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "to_dense before build");
  const int64_t lm = tree.lmax;
  auto d = download(D.get(), D.size(), stream);
  for (size_t i = 0; i < tree.leaves.size(); ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    for (int64_t cc = 0; cc < nd.len; ++cc)
      for (int64_t rr = 0; rr < nd.len; ++rr)
        A[(nd.start + rr) + (nd.start + cc) * lda] = d[i * lm * lm + rr + cc * lm];
  }
  for (int c = 1; c <= tree.max_depth; ++c) {
    HLevel& L = lv[c];
    auto us = download(L.US.get(), L.US.size(), stream);
    auto v = download(L.V.get(), L.V.size(), stream);
    for (size_t s = 0; s < L.ids.size(); ++s) {
      const TNode& rn = tree.nodes[L.ids[s]];
      const TNode& cn = tree.nodes[L.ids[L.sib[s]]];
      const int r = L.rank[s];
      for (int64_t j = 0; j < cn.len; ++j)
        for (int64_t i = 0; i < rn.len; ++i) {
          double acc = 0.0;
          for (int t = 0; t < r; ++t) acc += us[(rn.start + i) + t * N] * v[(cn.start + j) + t * N];
          A[(rn.start + i) + (cn.start + j) * lda] = acc;
        }
    }
  }
}

}  // namespace hmat

namespace hmat {

// Exchange (U S) and V at every depth, take ranks from the sibling,
// transpose the leaf blocks, and swap in the other factorisation.  Applying
// it twice restores the original state (sib is an involution).
void HodlrEngine::flip() {
  bg_.drop();  // the recorded batch graph writes the arrays swapped here
  for (int c = 1; c <= tree.max_depth; ++c) {
    HLevel& L = lv[c];
    std::swap(L.US, L.V);
    std::vector<int> r(L.rank.size());
    for (size_t s = 0; s < r.size(); ++s) r[s] = L.rank[L.sib[s]];
    L.rank.swap(r);
  }
  std::vector<int32_t> lens;
  for (int id : tree.leaves) lens.push_back((int32_t)tree.nodes[id].len);
  kern::transpose_blocks(D.get(), tree.lmax, lens, stream);
  std::swap(f64_, f64T_);
  std::swap(f32_, f32T_);
  std::swap(factored, factoredT_);
  transposed_ = !transposed_;
  prepare_matvec();
}

void HodlrEngine::matvec_transpose(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "matvec before build");
  flip();
  try {
    matvec(dX, ldx, nrhs, dY, ldy);
  } catch (...) {
    flip();
    throw;
  }
  flip();
}

void HodlrEngine::solve_transpose(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "solve before build");
  flip();
  try {
    solve(dB, ldb, nrhs, dX, ldx);
  } catch (...) {
    flip();
    throw;
  }
  flip();
}

}  // namespace hmat

namespace hmat {

// Operator file (little-endian, version 1):
//   magic "HMATHODLR\0" (10 bytes), int32 version
//   hmat_options opt, int64 N, int32 leaf, int32 nsys, int64 sizes[nsys]
//   int32 max_depth, int64 lmax, int32 nleaf
//   per depth c = 1..max_depth: int32 nseg, int32 rmax, int32 rank[nseg],
//     double US[N * rmax], double V[N * rmax]   (column-major, ld N)
//   double D[nleaf * lmax * lmax]
// The tree is rebuilt from (sizes, leaf); factorisations are not stored.
static const char kMagic[10] = {'H', 'M', 'A', 'T', 'H', 'O', 'D', 'L', 'R', 0};

void HodlrEngine::save(const std::string& path) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "save before build");
  HMAT_REQUIRE(!transposed_, HMAT_ERR_STATE, "save while transposed");
  std::ofstream f(path, std::ios::binary);
  HMAT_REQUIRE(f.good(), HMAT_ERR_ARG, "cannot open " + path + " for writing");
  auto put = [&](const void* p, size_t n) { f.write(reinterpret_cast<const char*>(p), (std::streamsize)n); };
  const int32_t version = 1;
  put(kMagic, sizeof(kMagic));
  put(&version, 4);
  put(&opt, sizeof(opt));
  put(&N, 8);
  const int32_t lf = leaf;
  put(&lf, 4);
  std::vector<int64_t> sizes;
  for (int id : tree.by_depth[0]) sizes.push_back(tree.nodes[id].len);
  const int32_t nsys = (int32_t)sizes.size();
  put(&nsys, 4);
  put(sizes.data(), 8 * sizes.size());
  const int32_t md = tree.max_depth;
  put(&md, 4);
  put(&tree.lmax, 8);
  const int32_t nl = (int32_t)tree.leaves.size();
  put(&nl, 4);
  for (int c = 1; c <= tree.max_depth; ++c) {
    const HLevel& L = lv[c];
    const int32_t ns = (int32_t)L.ids.size(), rm = L.rmax;
    put(&ns, 4);
    put(&rm, 4);
    std::vector<int32_t> rk(L.rank.begin(), L.rank.end());
    put(rk.data(), 4 * rk.size());
    auto us = download(L.US.get(), (size_t)N * rm, stream);
    auto v = download(L.V.get(), (size_t)N * rm, stream);
    put(us.data(), 8 * us.size());
    put(v.data(), 8 * v.size());
  }
  auto d = download(D.get(), (size_t)nl * tree.lmax * tree.lmax, stream);
  put(d.data(), 8 * d.size());
  HMAT_REQUIRE(f.good(), HMAT_ERR_ARG, "write to " + path + " failed");
}

void HodlrEngine::load(const std::string& path) {
  // Everything is read and checked into locals first; the handle's
  // operator, options and tree change only once the whole file is good.
  std::ifstream f(path, std::ios::binary);
  HMAT_REQUIRE(f.good(), HMAT_ERR_ARG, "cannot open " + path);
  auto get = [&](void* p, size_t n) {
    f.read(reinterpret_cast<char*>(p), (std::streamsize)n);
    HMAT_REQUIRE(f.good(), HMAT_ERR_ARG, path + ": truncated operator file");
  };
  char magic[10];
  int32_t version = 0;
  get(magic, 10);
  get(&version, 4);
  HMAT_REQUIRE(std::memcmp(magic, kMagic, 10) == 0 && version == 1, HMAT_ERR_ARG,
               path + ": not a version-1 HODLR operator file");
  hmat_options o;
  get(&o, sizeof(o));
  try {
    check_options(o);
  } catch (const Error& e) {
    throw Error(HMAT_ERR_ARG, path + ": bad options (" + e.what() + ")");
  }
  int64_t n = 0;
  int32_t lf = 0, nsys = 0;
  get(&n, 8);
  get(&lf, 4);
  get(&nsys, 4);
  HMAT_REQUIRE(n > 0 && lf > 0 && nsys > 0 && nsys <= n, HMAT_ERR_ARG, path + ": bad header");
  std::vector<int64_t> sizes(nsys);
  get(sizes.data(), 8 * (size_t)nsys);
  int64_t tot = 0;
  for (int64_t v : sizes) {
    HMAT_REQUIRE(v > 0 && v <= n, HMAT_ERR_ARG, path + ": bad system size");
    tot += v;
  }
  HMAT_REQUIRE(tot == n, HMAT_ERR_ARG, path + ": sizes do not add up to N");
  Tree t;
  if (nsys == 1)
    t.build(n, lf);
  else
    t.build_forest(sizes, lf);
  int32_t md = 0, nl = 0;
  int64_t lm = 0;
  get(&md, 4);
  get(&lm, 8);
  get(&nl, 4);
  HMAT_REQUIRE(md == t.max_depth && lm == t.lmax && nl == (int32_t)t.leaves.size(), HMAT_ERR_ARG,
               path + ": tree does not match the header");
  struct LevelData {
    std::vector<int> rank;
    int rmax = 0;
    std::vector<double> us, v;
  };
  std::vector<LevelData> ld(t.max_depth + 1);
  for (int c = 1; c <= t.max_depth; ++c) {
    const int ns = (int)t.by_depth[c].size();
    int32_t fns = 0, rm = 0;
    get(&fns, 4);
    get(&rm, 4);
    HMAT_REQUIRE(fns == ns && rm >= 0 && rm <= n, HMAT_ERR_ARG, path + ": level layout does not match the tree");
    std::vector<int32_t> rk(ns);
    get(rk.data(), 4 * (size_t)ns);
    for (int s = 0; s < ns; ++s) {
      // a block's rank is at most its smaller side
      const TNode& nd = t.nodes[t.by_depth[c][s]];
      const TNode& pp = t.nodes[nd.parent];
      const int64_t sib_len = t.nodes[(pp.left == t.by_depth[c][s]) ? pp.right : pp.left].len;
      HMAT_REQUIRE(rk[s] >= 0 && rk[s] <= rm && rk[s] <= std::min(nd.len, sib_len), HMAT_ERR_ARG,
                   path + ": bad rank");
    }
    ld[c].rank.assign(rk.begin(), rk.end());
    ld[c].rmax = rm;
    ld[c].us.resize((size_t)n * rm);
    ld[c].v.resize((size_t)n * rm);
    get(ld[c].us.data(), 8 * ld[c].us.size());
    get(ld[c].v.data(), 8 * ld[c].v.size());
  }
  std::vector<double> d((size_t)nl * lm * lm);
  get(d.data(), 8 * d.size());
  HMAT_REQUIRE(f.peek() == std::char_traits<char>::eof(), HMAT_ERR_ARG, path + ": trailing bytes after the operator");

  // commit: the file's options replace the handle's; a mixed-precision
  // operator refines to the default target unless one was set on this handle
  if (o.precision == HMAT_PREC_MIXED && !(refine_rtol > 0.0)) refine_rtol = 1e-10;
  opt = o;
  reset_for_build(lf);
  N = n;
  tree = std::move(t);
  lv.clear();
  lv.resize(tree.max_depth + 1);
  par.clear();
  for (int c = 1; c <= tree.max_depth; ++c) {
    HLevel& L = lv[c];
    L.ids = tree.by_depth[c];
    const int ns = (int)L.ids.size();
    L.sib.assign(ns, 0);
    for (int s = 0; s < ns; ++s) {
      const TNode& nd = tree.nodes[L.ids[s]];
      const TNode& par_ = tree.nodes[nd.parent];
      const int sib_id = (par_.left == L.ids[s]) ? par_.right : par_.left;
      L.sib[s] = tree.nodes[sib_id].pos;
    }
    L.rank = std::move(ld[c].rank);
    L.rmax = ld[c].rmax;
    L.US.upload(ld[c].us, stream);
    L.V.upload(ld[c].v, stream);
    HIP_CHECK(hipStreamSynchronize(stream));
    L.S.assign((size_t)ns * L.rmax, 0.0);
  }
  D.upload(d, stream);
  HIP_CHECK(hipStreamSynchronize(stream));
  prepare_matvec();
  built = true;
}

}  // namespace hmat
