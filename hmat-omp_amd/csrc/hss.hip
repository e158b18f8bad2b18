// HSS engine on gfx950.
//
// Reference: HSSMatrixOMP (include/hss_routines_omp.h).  The reference
// builds leaf bases from whole off-diagonal row/column blocks, keeps
// identity transfers (so the root rank is the sum of all leaf ranks), forms
// every coupling from dense reconstructed bases, and solves column by column
// through a recursion that re-solves the children once per basis column.
// Here:
//   * leaf bases: all leaves in one batch (generated sketch that excludes the
//     leaf's own columns, CholeskyQR, Jacobi SVD of k x k) — with the same
//     Omega the bases are the reference's;
//   * transfers and couplings from the projected blocks
//     T_p = U_p^T A(I_p, I_p^c): one pass over A per side, none per level;
//   * matvec: up-sweep g = W^T [g_a; g_b], down-sweep f = R f + B g
//     (hss_routines_omp.h:154-230);
//   * solve: the reference's push-through Woodbury (hss_routines_omp.h:
//     306-343) restated on nested bases, Q_p = A_p^{-1} U_p = diag(Q_a, Q_b) G_p,
//     factored once bottom-up and applied by one up-sweep and one down-sweep.
#include <algorithm>
#include <fstream>
#include <cmath>
#include <cstring>

#include "hss.hpp"
#include "sweep.hpp"

namespace hmat {

using kern::BGemm;
using kern::Rect;
using kern::ZeroRect;

HssEngine::HssEngine(const hmat_options& o) : opt(o) {
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  own_stream = true;
}
HssEngine::~HssEngine() {
  if (stream) stager_release(stream);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

// Reverse node id over leaves, row block then column block (hss_routines.h:57-77).
int64_t HssEngine::omega_length(const hmat_options& o, int64_t N, int leaf) {
  Tree t;
  t.build(N, leaf);
  if (t.max_depth == 0) return 0;
  int64_t tot = 0;
  for (int id : t.leaves) {
    const int64_t m = t.nodes[id].len;
    if (uses_rsvd(m, N, o)) tot += 2 * N * sketch_width(m, N, o);
  }
  return tot;
}

int HssEngine::root_rank() const {
  if (!built || tree.max_depth == 0) return 0;
  const TNode& r = tree.nodes[0];
  return sn[r.left].rU + sn[r.right].rU;
}

int64_t HssEngine::pool_alloc(int64_t n) {
  const int64_t off = pool_used;
  pool_used += round_up((int)std::max<int64_t>(n, 1), 32);
  if ((size_t)pool_used > pool.size()) {
    DevBuf<double> bigger((size_t)std::max<int64_t>(pool_used * 2, 1 << 16));
    if (off > 0)
      HIP_CHECK(hipMemcpyAsync(bigger.get(), pool.get(), off * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    pool = std::move(bigger);
  }
  return off;
}

void HssEngine::leaf_apply(const double* Dm, const double* B, int64_t ldb, int ncols, double* Y,
                           int64_t ldy, double alpha, double beta) {
  const int64_t lm = tree.lmax;
  std::vector<SegNN> segs;
  for (size_t i = 0; i < tree.leaves.size(); ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    segs.push_back(SegNN{(int64_t)i * lm * lm, nd.start, nd.start, (int32_t)nd.len, (int32_t)nd.len, ncols, 0});
  }
  kern::seg_gemm_nn<double>(segs, Dm, lm, B, ldb, Y, ldy, alpha, beta, stream);
}

void HssEngine::leaf_solve(const double* B, int64_t ldb, int ncols, double* Y, int64_t ldy, double* R) {
  leaf_apply(Dinv.get(), B, ldb, ncols, Y, ldy);
  kern::convert_cols<double, double>(N, ncols, B, ldb, R, N, stream);
  leaf_apply(D.get(), Y, ldy, ncols, R, N, -1.0, 1.0);
  leaf_apply(Dinv.get(), R, N, ncols, Y, ldy, 1.0, 1.0);
}

// Leaf bases of one side.  Row side: range of A(I, I^c) (len x N with the
// leaf's own columns excluded) -> U_i, and T_i^T = A(I, I^c)^T U_i (N x r).
// Column side: range of A(I^c, I)^T -> V_i and S_i^T = A(I^c, I) V_i.
// (hss_routines_omp.h:77-93; compression.h:67-99)
void HssEngine::compress_leaves(const GenDesc& g, bool col, const double* host_omega,
                                const std::vector<int64_t>& omega_off) {
  const auto& L = tree.leaves;
  const int nl = (int)L.size();
  std::vector<BlockMode> mode(nl);
  std::vector<int> k(nl);
  int kcap = 0;
  for (int i = 0; i < nl; ++i) {
    const int64_t m = tree.nodes[L[i]].len;
    if (uses_rsvd(m, N, opt)) {
      mode[i] = BlockMode::RSVD;
      k[i] = sketch_width(m, N, opt);
    } else {
      mode[i] = BlockMode::ExactWide;  // m < N: Q = I_m
      k[i] = (int)m;
    }
    kcap = std::max(kcap, k[i]);
  }
  HMAT_REQUIRE(kcap <= 300, HMAT_ERR_ARG, "leaf sketch width above 300 is not supported");
  const int64_t nk = N * (int64_t)kcap;
  LeafScratch& sc = ls_;
  DevBuf<double>&Yb = sc.Yb, &Tb = sc.Tb, &Bw = sc.Bw, &Tw2 = sc.Tw2;
  Yb.reserve(nk);
  Tb.reserve(nk);
  Bw.reserve((size_t)nk * nl);
  Tw2.reserve((size_t)nk * nl);
  Yb.zero(stream);
  Bw.zero(stream);
  // Omega: per leaf (host, reference draw order) or one shared device draw
  const bool per_leaf = host_omega != nullptr;
  DevBuf<double>& X = sc.X;
  X.reserve(per_leaf ? (size_t)nk * nl : (size_t)nk);
  X.zero(stream);
  if (per_leaf) {
    std::vector<double> h((size_t)nk * nl, 0.0);
    for (int i = 0; i < nl; ++i) {
      if (mode[i] != BlockMode::RSVD) continue;
      const int64_t o = omega_off[2 * L[i] + (col ? 1 : 0)];
      HMAT_REQUIRE(o >= 0, HMAT_ERR_STATE, "omega schedule mismatch");
      std::memcpy(&h[(size_t)i * nk], host_omega + o, sizeof(double) * N * k[i]);
    }
    HIP_CHECK(hipMemcpyAsync(X.get(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  } else {
    std::vector<SegNN> rnd{SegNN{0, col ? 0x77777 : 0x55555, 0, (int32_t)N, 0, kcap, 0}};
    kern::fill_random_segs(rnd, dnn_, X.get(), N, opt.seed, opt.omega_dist, stream);
  }
  std::vector<SegNN> eye;
  std::vector<SegSk> fwd, fwd_r, bwd, bwd_r;
  std::vector<OrthSeg> oy, ob, ob_r;
  std::vector<ZeroRect> zr;
  for (int i = 0; i < nl; ++i) {
    const TNode& nd = tree.nodes[L[i]];
    const int64_t s = nd.start, m = nd.len;
    const int64_t xo = per_leaf ? (int64_t)i * nk : 0;
    const int64_t bo = (int64_t)i * nk;
    // forward: Y(rows I) = A(I, I^c) X   (column side: A(I^c, I)^T X)
    const SegSk f{s, 0, s, xo, s, (int32_t)m, (int32_t)N, (int32_t)m, k[i]};
    const SegSk fz{s, 0, s, bo, s, (int32_t)m, (int32_t)N, (int32_t)m, k[i]};
    // backward: B(all rows) = A(I, :)^T Q   (column side: A(:, I) Q), own rows zeroed after
    const SegSk b{0, s, 0, s, bo, (int32_t)N, (int32_t)m, 0, k[i]};
    bwd.push_back(b);
    ob.push_back(OrthSeg{bo, (int32_t)N, k[i]});
    zr.push_back(ZeroRect{Bw.get() + bo, N, (int32_t)s, (int32_t)m, 0, k[i]});
    if (mode[i] == BlockMode::RSVD) {
      fwd.push_back(f);
      fwd_r.push_back(fz);
      bwd_r.push_back(b);
      oy.push_back(OrthSeg{s, (int32_t)m, k[i]});
      ob_r.push_back(OrthSeg{bo, (int32_t)N, k[i]});
    } else {
      eye.push_back(SegNN{0, 0, s, (int32_t)m, 0, k[i], 0});
    }
  }
  const bool tr_fwd = col;  // row side reads A(I, red); column side A(red, I)^T
  kern::fill_identity_segs(eye, dnn_, Yb.get(), N, stream);
  kern::sketch(g, tr_fwd, fwd, X.get(), N, Yb.get(), N, kcap, ws_, stream);
  for (int it = 0; it < opt.power_iters && !fwd_r.empty(); ++it) {
    orth_.run(oy, Yb.get(), Tb.get(), N, kcap, nullptr, stream);
    kern::sketch(g, !tr_fwd, bwd_r, Yb.get(), N, Bw.get(), N, kcap, ws_, stream);
    kern::zero_rects(zr, dzr_, stream);
    orth_.run(ob_r, Bw.get(), Tw2.get(), N, kcap, nullptr, stream);
    kern::sketch(g, tr_fwd, fwd_r, Bw.get(), N, Yb.get(), N, kcap, ws_, stream);
  }
  orth_.run(oy, Yb.get(), Tb.get(), N, kcap, nullptr, stream);
  kern::sketch(g, !tr_fwd, bwd, Yb.get(), N, Bw.get(), N, kcap, ws_, stream);
  kern::zero_rects(zr, dzr_, stream);
  const int64_t kk = (int64_t)kcap * kcap;
  DevBuf<double>&Rb = sc.Rb, &Vr = sc.Vr, &Sv = sc.Sv, &Wk = sc.Wk, &TU = sc.TU, &TV = sc.TV;
  Rb.reserve((size_t)nl * kk);
  Vr.reserve((size_t)nl * kk);
  Sv.reserve((size_t)nl * kcap);
  Wk.reserve((size_t)2 * nl * kk);
  TU.reserve((size_t)nl * kk);
  TV.reserve((size_t)nl * kk);
  orth_.run(ob, Bw.get(), Tw2.get(), N, kcap, Rb.get(), stream);
  std::vector<SmallMat> sm(nl);
  for (int i = 0; i < nl; ++i) sm[i] = SmallMat{(int64_t)i * kk, k[i], kcap};
  kern::jacobi_svd(sm, dsm_, Rb.get(), Vr.get(), Sv.get(), kcap, Wk.get(), stream);
  auto S = download(Sv.get(), (size_t)nl * kcap, stream);
  std::vector<int> rank(nl);
  int rmax = 0;
  for (int i = 0; i < nl; ++i) {
    const double* sv = &S[(size_t)i * kcap];
    for (int j = 0; j < k[i]; ++j)
      if (!std::isfinite(sv[j])) throw Error(HMAT_ERR_NONFINITE, "non-finite singular value in HSS build");
    int r = 0;
    if (opt.fixed_rank) {
      r = std::min(opt.rank, k[i]);
    } else {
      if (sv[0] > 1e-15)
        for (int j = 0; j < k[i]; ++j) r += (sv[j] > opt.tol * sv[0]) ? 1 : 0;
      if (r == 0) r = 1;
    }
    rank[i] = r;
    rmax = std::max(rmax, r);
  }
  DevBuf<int> drank;
  drank.upload(rank, stream);
  // basis = Q Vr(:, :r); projected block^T = Qb Ur(:, :r) S(:r)
  kern::trunc_prep(sm, dsm_, Rb.get(), Vr.get(), Sv.get(), kcap, drank.get(), rmax, TU.get(), TV.get(), 1,
                   stream);
  DevBuf<double>& basis = col ? Vl : Ul;
  basis.resize((size_t)N * rmax);
  basis.zero(stream);
  std::vector<int64_t>& tcol = leafcol[col ? 1 : 0];
  tcol.assign(nl, 0);
  int64_t tsum = 0;
  for (int i = 0; i < nl; ++i) {
    tcol[i] = tsum;
    tsum += rank[i];
  }
  DevBuf<double>& Tleaf = leafT[col ? 1 : 0];
  Tleaf.reserve((size_t)N * std::max<int64_t>(tsum, 1));
  std::vector<SegNN> nb, nt;
  for (int i = 0; i < nl; ++i) {
    const TNode& nd = tree.nodes[L[i]];
    nb.push_back(SegNN{nd.start, (int64_t)i * kk, nd.start, (int32_t)nd.len, k[i], rank[i], 0});
    nt.push_back(SegNN{(int64_t)i * nk, (int64_t)i * kk, tcol[i] * N, (int32_t)N, k[i], rank[i], 0});
  }
  kern::seg_gemm_nn(nb, Yb.get(), N, TU.get(), kcap, basis.get(), N, 1.0, 0.0, stream);
  kern::seg_gemm_nn(nt, Bw.get(), N, TV.get(), kcap, Tleaf.get(), N, 1.0, 0.0, stream);
  HIP_CHECK(hipStreamSynchronize(stream));
  for (int i = 0; i < nl; ++i) {
    if (col) sn[L[i]].rV = rank[i];
    else sn[L[i]].rU = rank[i];
  }
  if (col) rVmax = rmax;
  else rUmax = rmax;
}

// Transfers (nested) or identities (reference), couplings and full column
// bases, depth by depth from the leaves up.
void HssEngine::build_internal() {
  const int dmax = tree.max_depth;
  const int nn = (int)tree.nodes.size();
  std::vector<int64_t> tcol(nn, 0), scol(nn, 0);
  // child depth: wide T^T, S^T and full V; the current depth's go to the
  // other buffer of each pair and the pairs swap after every depth
  NodeScratch& sc = ns_;
  DevBuf<double>*Tc = &sc.T[0], *Sc = &sc.S[0], *VFc = &sc.VF[0];
  DevBuf<double>*Tn = &sc.T[1], *Sn = &sc.S[1], *VFn = &sc.VF[1];

  // wide projection arrays of one depth: leaves from the leaf arrays,
  // internal nodes from (Tin, Sin) at columns (icT, icS)
  auto make_wide = [&](const std::vector<int>& ids, const double* Tin, const double* Sin,
                       const std::vector<int64_t>& icT, const std::vector<int64_t>& icS, DevBuf<double>& Tw,
                       DevBuf<double>& Sw) {
    int64_t ts = 0, ss = 0;
    for (int id : ids) {
      tcol[id] = ts;
      ts += sn[id].rU;
      scol[id] = ss;
      ss += sn[id].rV;
    }
    Tw.reserve((size_t)N * std::max<int64_t>(ts, 1));
    Sw.reserve((size_t)N * std::max<int64_t>(ss, 1));
    std::vector<SegNN> lt, ls, it, is;
    for (int id : ids) {
      const TNode& nd = tree.nodes[id];
      if (nd.leaf) {
        lt.push_back(SegNN{leafcol[0][nd.leaf_idx] * N, 0, tcol[id] * N, (int32_t)N, 0, sn[id].rU, 0});
        ls.push_back(SegNN{leafcol[1][nd.leaf_idx] * N, 0, scol[id] * N, (int32_t)N, 0, sn[id].rV, 0});
      } else {
        it.push_back(SegNN{icT[id] * N, 0, tcol[id] * N, (int32_t)N, 0, sn[id].rU, 0});
        is.push_back(SegNN{icS[id] * N, 0, scol[id] * N, (int32_t)N, 0, sn[id].rV, 0});
      }
    }
    kern::copy_segs(lt, dnn_, leafT[0].get(), N, Tw.get(), N, stream);
    kern::copy_segs(ls, dnn_, leafT[1].get(), N, Sw.get(), N, stream);
    if (!it.empty()) {
      kern::copy_segs(it, dnn_, Tin, N, Tw.get(), N, stream);
      kern::copy_segs(is, dnn_, Sin, N, Sw.get(), N, stream);
    }
  };
  // full column bases of one depth: V_leaf, or diag(V_a, V_b) W_p
  auto make_vf = [&](const std::vector<int>& ids, DevBuf<double>& VF) {
    int rv = 1, ldw = 1;
    for (int id : ids) {
      rv = std::max(rv, sn[id].rV);
      if (!tree.nodes[id].leaf) ldw = sn[id].ldW;
    }
    VF.reserve((size_t)N * rv);
    VF.zero(stream);
    std::vector<SegNN> cp, mm;
    for (int id : ids) {
      const TNode& nd = tree.nodes[id];
      if (nd.leaf) {
        cp.push_back(SegNN{nd.start, 0, nd.start, (int32_t)nd.len, 0, sn[id].rV, 0});
      } else {
        const TNode& a = tree.nodes[nd.left];
        const TNode& b = tree.nodes[nd.right];
        const int va = sn[nd.left].rV;
        mm.push_back(SegNN{a.start, sn[id].oW, a.start, (int32_t)a.len, va, sn[id].rV, 0});
        mm.push_back(SegNN{b.start, sn[id].oW + va, b.start, (int32_t)b.len, sn[nd.right].rV, sn[id].rV, 0});
      }
    }
    kern::copy_segs(cp, dnn_, Vl.get(), N, VF.get(), N, stream);
    kern::seg_gemm_nn(mm, VFc->get(), N, pool.get(), ldw, VF.get(), N, 1.0, 0.0, stream);
  };

  const std::vector<int64_t> none(nn, 0);
  make_wide(tree.by_depth[dmax], nullptr, nullptr, none, none, *Tc, *Sc);
  make_vf(tree.by_depth[dmax], *VFc);

  for (int d = dmax - 1; d >= 0; --d) {
    const auto& ids = tree.by_depth[d];
    std::vector<int> inner;
    for (int id : ids)
      if (!tree.nodes[id].leaf) inner.push_back(id);
    // ---- couplings B_ab = T_a(:, I_b) V_b^full, B_ba = T_b(:, I_a) V_a^full
    // (B12 = U_L^T A12 V_R, hss_routines_omp.h:137-143)
    int ldb = 1;
    for (int id : inner)
      ldb = std::max(ldb, std::max(sn[tree.nodes[id].left].rU, sn[tree.nodes[id].right].rU));
    for (int id : inner) {
      const TNode& nd = tree.nodes[id];
      sn[id].ldB = ldb;
      sn[id].oBab = pool_alloc((int64_t)ldb * sn[nd.right].rV);
      sn[id].oBba = pool_alloc((int64_t)ldb * sn[nd.left].rV);
    }
    {
      std::vector<SegTN> tb;
      for (int id : inner) {
        const TNode& nd = tree.nodes[id];
        const TNode& a = tree.nodes[nd.left];
        const TNode& b = tree.nodes[nd.right];
        tb.push_back(SegTN{tcol[nd.left] * N + b.start, b.start, sn[id].oBab, (int32_t)b.len, sn[nd.left].rU,
                           sn[nd.right].rV, 0});
        tb.push_back(SegTN{tcol[nd.right] * N + a.start, a.start, sn[id].oBba, (int32_t)a.len, sn[nd.right].rU,
                           sn[nd.left].rV, 0});
      }
      kern::seg_gemm_tn(tb, Tc->get(), N, VFc->get(), N, pool.get(), ldb, 1.0, 0.0, ws_, stream);
    }
    if (d == 0) break;  // the root has no transfer and no projection
    // ---- stacked child projections P^T = [T_a^T | T_b^T], own rows zeroed
    std::vector<int> kR(nn, 0), kW(nn, 0);
    int kcap = 1, ldR = 1, ldW = 1;
    int64_t psum = 0, qsum = 0;
    std::vector<int64_t> pcol(nn, 0), qcol(nn, 0);
    for (int id : inner) {
      const TNode& nd = tree.nodes[id];
      kR[id] = sn[nd.left].rU + sn[nd.right].rU;
      kW[id] = sn[nd.left].rV + sn[nd.right].rV;
      kcap = std::max(kcap, std::max(kR[id], kW[id]));
      ldR = std::max(ldR, kR[id]);
      ldW = std::max(ldW, kW[id]);
      pcol[id] = psum;
      psum += kR[id];
      qcol[id] = qsum;
      qsum += kW[id];
    }
    DevBuf<double>&Pt = sc.Pt, &Ps = sc.Ps;
    Pt.reserve((size_t)N * std::max<int64_t>(psum, 1));
    Ps.reserve((size_t)N * std::max<int64_t>(qsum, 1));
    {
      std::vector<SegNN> ct, cs;
      std::vector<ZeroRect> zr;
      for (int id : inner) {
        const TNode& nd = tree.nodes[id];
        const int ua = sn[nd.left].rU, va = sn[nd.left].rV;
        ct.push_back(SegNN{tcol[nd.left] * N, 0, pcol[id] * N, (int32_t)N, 0, ua, 0});
        ct.push_back(SegNN{tcol[nd.right] * N, 0, (pcol[id] + ua) * N, (int32_t)N, 0, sn[nd.right].rU, 0});
        cs.push_back(SegNN{scol[nd.left] * N, 0, qcol[id] * N, (int32_t)N, 0, va, 0});
        cs.push_back(SegNN{scol[nd.right] * N, 0, (qcol[id] + va) * N, (int32_t)N, 0, sn[nd.right].rV, 0});
        zr.push_back(ZeroRect{Pt.get() + pcol[id] * N, N, (int32_t)nd.start, (int32_t)nd.len, 0, kR[id]});
        zr.push_back(ZeroRect{Ps.get() + qcol[id] * N, N, (int32_t)nd.start, (int32_t)nd.len, 0, kW[id]});
      }
      kern::copy_segs(ct, dnn_, Tc->get(), N, Pt.get(), N, stream);
      kern::copy_segs(cs, dnn_, Sc->get(), N, Ps.get(), N, stream);
      kern::zero_rects(zr, dzr_, stream);
    }
    std::vector<int64_t> ncT(nn, 0), ncS(nn, 0);
    if (!nested) {
      // identity transfers, ranks add up (hss_routines_omp.h:110-119)
      std::vector<Rect> eye;
      for (int id : inner) {
        sn[id].ldR = ldR;
        sn[id].ldW = ldW;
        sn[id].oR = pool_alloc((int64_t)ldR * kR[id]);
        sn[id].oW = pool_alloc((int64_t)ldW * kW[id]);
      }
      for (int id : inner) {
        eye.push_back(Rect{nullptr, pool.get() + sn[id].oR, 0, ldR, kR[id], kR[id], 1.0, 0.0});
        eye.push_back(Rect{nullptr, pool.get() + sn[id].oW, 0, ldW, kW[id], kW[id], 1.0, 0.0});
      }
      kern::rect_axpby(eye, drc_, stream);  // zero
      kern::rect_add_identity(eye, drc_, stream);
      for (int id : inner) {
        sn[id].rU = kR[id];
        sn[id].rV = kW[id];
        ncT[id] = pcol[id];
        ncS[id] = qcol[id];
      }
      make_wide(ids, Pt.get(), Ps.get(), ncT, ncS, *Tn, *Sn);
    } else {
      // compress P^T (N x k) per node: P^T = Qp R, R = Ur S Vr^T ->
      // transfer Vr(:, :r), new projection^T = Qp Ur(:, :r) S(:r)
      const int ni = (int)inner.size();
      const int64_t kk = (int64_t)kcap * kcap;
      DevBuf<double>* Pn = sc.Pn;
      for (int side = 0; side < 2; ++side) {
        DevBuf<double>& P = side ? Ps : Pt;
        const std::vector<int64_t>& pc = side ? qcol : pcol;
        const std::vector<int>& kx = side ? kW : kR;
        const int ldx = side ? ldW : ldR;
        DevBuf<double>&Tmp = sc.Tmp, &Rb = sc.Rb, &Vr = sc.Vr, &Sv = sc.Sv, &Wk = sc.Wk, &TU = sc.TU, &TV = sc.TV;
        Tmp.reserve(P.size());
        Rb.reserve((size_t)ni * kk);
        Vr.reserve((size_t)ni * kk);
        Sv.reserve((size_t)ni * kcap);
        Wk.reserve((size_t)2 * ni * kk);
        TU.reserve((size_t)ni * kk);
        TV.reserve((size_t)ni * kk);
        std::vector<OrthSeg> os;
        std::vector<SmallMat> sm;
        for (int i = 0; i < ni; ++i) {
          os.push_back(OrthSeg{pc[inner[i]] * N, (int32_t)N, kx[inner[i]]});
          sm.push_back(SmallMat{(int64_t)i * kk, kx[inner[i]], kcap});
        }
        orth_.run(os, P.get(), Tmp.get(), N, kcap, Rb.get(), stream);
        kern::jacobi_svd(sm, dsm_, Rb.get(), Vr.get(), Sv.get(), kcap, Wk.get(), stream);
        auto S = download(Sv.get(), (size_t)ni * kcap, stream);
        std::vector<int> rk(ni);
        int rm = 1;
        for (int i = 0; i < ni; ++i) {
          const double* sv = &S[(size_t)i * kcap];
          const int kxi = kx[inner[i]];
          int r = 0;
          if (opt.fixed_rank) {
            r = std::min(opt.rank, kxi);
          } else {
            if (sv[0] > 1e-15)
              for (int j = 0; j < kxi; ++j) r += (sv[j] > opt.tol * sv[0]) ? 1 : 0;
            if (r == 0) r = 1;
          }
          rk[i] = r;
          rm = std::max(rm, r);
        }
        DevBuf<int> drk;
        drk.upload(rk, stream);
        kern::trunc_prep(sm, dsm_, Rb.get(), Vr.get(), Sv.get(), kcap, drk.get(), rm, TU.get(), TV.get(), 1,
                         stream);
        for (int i = 0; i < ni; ++i) {
          const int id = inner[i];
          if (side) {
            sn[id].ldW = ldx;
            sn[id].oW = pool_alloc((int64_t)ldx * rk[i]);
          } else {
            sn[id].ldR = ldx;
            sn[id].oR = pool_alloc((int64_t)ldx * rk[i]);
          }
        }
        std::vector<Rect> cp;
        int64_t sum = 0;
        std::vector<SegNN> nt;
        std::vector<int64_t>& nc = side ? ncS : ncT;
        for (int i = 0; i < ni; ++i) {
          const int id = inner[i];
          const int64_t o = side ? sn[id].oW : sn[id].oR;
          cp.push_back(Rect{TU.get() + (int64_t)i * kk, pool.get() + o, kcap, ldx, kx[id], rk[i], 1.0, 0.0});
          nc[id] = sum;
          nt.push_back(SegNN{pc[id] * N, (int64_t)i * kk, sum * N, (int32_t)N, kx[id], rk[i], 0});
          sum += rk[i];
        }
        kern::rect_axpby(cp, drc_, stream);
        Pn[side].reserve((size_t)N * std::max<int64_t>(sum, 1));
        kern::seg_gemm_nn(nt, P.get(), N, TV.get(), kcap, Pn[side].get(), N, 1.0, 0.0, stream);
        for (int i = 0; i < ni; ++i) (side ? sn[inner[i]].rV : sn[inner[i]].rU) = rk[i];
      }
      make_wide(ids, Pn[0].get(), Pn[1].get(), ncT, ncS, *Tn, *Sn);
    }
    make_vf(ids, *VFn);
    std::swap(Tc, Tn);
    std::swap(Sc, Sn);
    std::swap(VFc, VFn);
  }
  HIP_CHECK(hipStreamSynchronize(stream));
}

int HssEngine::sampled_width(const hmat_options& o, int64_t N) {
  return (int)std::min<int64_t>((int64_t)o.rank + o.oversample, N);
}
int64_t HssEngine::sampled_omega_length(const hmat_options& o, int64_t N) {
  return 2 * N * (int64_t)sampled_width(o, N);
}

int64_t HssEngine::spool_alloc(int64_t n) {
  SampleScratch& sc = ss_;
  const int64_t off = sc.sp_used;
  sc.sp_used += round_up((int)std::max<int64_t>(n, 1), 32);
  if ((size_t)sc.sp_used > sc.sp.size()) {
    DevBuf<double> bigger((size_t)std::max<int64_t>(sc.sp_used * 2, 1 << 16));
    if (off > 0)
      HIP_CHECK(hipMemcpyAsync(bigger.get(), sc.sp.get(), off * sizeof(double), hipMemcpyDeviceToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    sc.sp = std::move(bigger);
  }
  return off;
}

// Nested generators from one sample pass per side (Martinsson 2011, SIMAX
// 32(4), orthonormal bases).  With s = rank + oversample Gaussian columns:
//   leaf i:   Yh_i = A(I, I^c) Omega(I^c)  (the sketch with the leaf's own
//             columns excluded), U_i = leading left singular vectors of Yh_i,
//             yh_i = U_i^T Yh_i; om_i = V_i^T Omega(I), ps_i = U_i^T Psi(I);
//             the column side likewise from Zh_i = A(I^c, I)^T Psi(I^c);
//   siblings: B_ab = U_a^T A(I_a, I_b) V_b (full bases; one generated-entry
//             product per depth, then a segmented U^T P);
//   parent p: M_p = [yh_a - B_ab om_b; yh_b - B_ba om_a] -> R_p (its leading
//             left singular vectors), yh_p = R_p^T M_p, om_p = W_p^T [om_a; om_b],
//             and the mirror image for W_p, zh_p, ps_p.
// The reduced samples are kept transposed (s rows) in one device pool:
// yhT, zhT, omT, psT of node p at spo[4p + 0..3], leading dimension s.
// Singular vectors of a small M (m x s) come from one-sided Jacobi on M^T
// zero-padded to max(m, s) square: M^T V = U S gives M = V S U^T, so the
// basis is V(:m, :r) and (the reduced samples)^T = U(:s, :r) S.  Tall leaf
// samples (len > s) are first factored Yh = Q R (CholeskyQR), Jacobi then
// runs on R^T and the basis is Q V(:, :r).
void HssEngine::build_sampled(const GenDesc& g, const double* host_omega, int64_t omega_len) {
  (void)omega_len;
  const int ks = sampled_width(opt, N);
  const auto& L = tree.leaves;
  const int nl = (int)L.size();
  const int nn = (int)tree.nodes.size();
  const int dmax = tree.max_depth;
  SampleScratch& sc = ss_;
  sc.sp_used = 0;
  std::vector<int64_t> spo(4 * (size_t)nn, -1);
  const int64_t nks = N * (int64_t)ks;
  sc.Om.reserve(nks);
  sc.Ps.reserve(nks);
  sc.Y.reserve(nks);
  sc.Z.reserve(nks);
  sc.Tb.reserve(nks);
  if (host_omega) {
    HIP_CHECK(hipMemcpyAsync(sc.Om.get(), host_omega, nks * sizeof(double), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(sc.Ps.get(), host_omega + nks, nks * sizeof(double), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  } else {
    std::vector<SegNN> r0{SegNN{0, 0x5A301, 0, (int32_t)N, 0, ks, 0}};
    std::vector<SegNN> r1{SegNN{0, 0x5A302, 0, (int32_t)N, 0, ks, 0}};
    kern::fill_random_segs(r0, dnn_, sc.Om.get(), N, opt.seed, opt.omega_dist, stream);
    kern::fill_random_segs(r1, dnn_, sc.Ps.get(), N, opt.seed, opt.omega_dist, stream);
  }
  // ---- leaf samples: Y(I) = A(I, I^c) Omega, Z(I) = A(I^c, I)^T Psi
  {
    std::vector<SegSk> sk;
    for (int id : L) {
      const TNode& nd = tree.nodes[id];
      sk.push_back(SegSk{nd.start, 0, nd.start, 0, nd.start, (int32_t)nd.len, (int32_t)N, (int32_t)nd.len, ks});
    }
    kern::sketch(g, false, sk, sc.Om.get(), N, sc.Y.get(), N, ks, ws_, stream);
    kern::sketch(g, true, sk, sc.Ps.get(), N, sc.Z.get(), N, ks, ws_, stream);
  }
  // rank rule of compression.h:42-46 on the singular values of a sample
  auto pick_rank = [&](const double* sv, int nsv) {
    for (int j = 0; j < nsv; ++j)
      if (!std::isfinite(sv[j])) throw Error(HMAT_ERR_NONFINITE, "non-finite singular value in HSS build");
    if (opt.fixed_rank) return std::max(1, std::min(opt.rank, nsv));
    int r = 0;
    if (sv[0] > 1e-15)
      for (int j = 0; j < nsv; ++j) r += (sv[j] > opt.tol * sv[0]) ? 1 : 0;
    return std::max(r, 1);
  };
  // ---- leaf bases, both sides
  const int64_t kk = (int64_t)ks * ks;
  sc.Sm.reserve((size_t)nl * kk);
  sc.Rt.reserve((size_t)nl * kk);
  sc.Vr.reserve((size_t)nl * kk);
  sc.Sv.reserve((size_t)nl * ks);
  sc.Wk.reserve((size_t)2 * nl * kk);
  sc.TU.reserve((size_t)nl * kk);
  sc.TV.reserve((size_t)nl * kk);
  for (int side = 0; side < 2; ++side) {
    double* Ys = side ? sc.Z.get() : sc.Y.get();
    std::vector<OrthSeg> tall;
    std::vector<int> tall_leaf;
    std::vector<Rect> small;
    for (int i = 0; i < nl; ++i) {
      const TNode& nd = tree.nodes[L[i]];
      if (nd.len > ks) {
        tall.push_back(OrthSeg{nd.start, (int32_t)nd.len, ks});
        tall_leaf.push_back(i);
      } else {
        small.push_back(Rect{Ys + nd.start, sc.Sm.get() + (int64_t)i * kk, N, ks, (int32_t)nd.len, ks, 1.0, 0.0});
      }
    }
    if (!small.empty()) zero_bytes_async(sc.Sm.get(), (size_t)nl * kk * sizeof(double), stream);
    if (!tall.empty()) {
      // Yh = Q R in place (Q overwrites the samples), then R^T into the leaf's slot
      orth_.run(tall, Ys, sc.Tb.get(), N, ks, sc.Rt.get(), stream);
      std::vector<Rect> rt;
      for (size_t b = 0; b < tall.size(); ++b)
        rt.push_back(Rect{sc.Rt.get() + (int64_t)b * kk, sc.Sm.get() + (int64_t)tall_leaf[b] * kk, ks, ks, ks, ks,
                          1.0, 0.0});
      kern::rect_transpose(rt, drc_, stream);
    }
    kern::rect_transpose(small, drc_, stream);
    std::vector<SmallMat> sm(nl);
    for (int i = 0; i < nl; ++i) sm[i] = SmallMat{(int64_t)i * kk, ks, ks};
    kern::jacobi_svd(sm, dsm_, sc.Sm.get(), sc.Vr.get(), sc.Sv.get(), ks, sc.Wk.get(), stream);
    const auto S = download(sc.Sv.get(), (size_t)nl * ks, stream);
    std::vector<int> rank(nl);
    int rmax = 1;
    for (int i = 0; i < nl; ++i) {
      const int nsv = (int)std::min<int64_t>(tree.nodes[L[i]].len, ks);
      rank[i] = pick_rank(&S[(size_t)i * ks], nsv);
      rmax = std::max(rmax, rank[i]);
    }
    DevBuf<int> drank;
    drank.upload(rank, stream);
    // TU = V(:, :r) (basis coefficients), TV = U(:, :r) S (reduced samples^T)
    kern::trunc_prep(sm, dsm_, sc.Sm.get(), sc.Vr.get(), sc.Sv.get(), ks, drank.get(), rmax, sc.TU.get(),
                     sc.TV.get(), 1, stream);
    DevBuf<double>& basis = side ? Vl : Ul;
    basis.resize((size_t)N * rmax);
    basis.zero(stream);
    std::vector<SegNN> qt;
    std::vector<Rect> cp, red;
    for (int i = 0; i < nl; ++i) {
      const int id = L[i];
      const TNode& nd = tree.nodes[id];
      if (nd.len > ks)
        qt.push_back(SegNN{nd.start, (int64_t)i * kk, nd.start, (int32_t)nd.len, ks, rank[i], 0});
      else
        cp.push_back(Rect{sc.TU.get() + (int64_t)i * kk, basis.get() + nd.start, ks, N, (int32_t)nd.len, rank[i],
                          1.0, 0.0});
      spo[4 * id + side] = spool_alloc((int64_t)ks * rank[i]);
      if (side) sn[id].rV = rank[i];
      else sn[id].rU = rank[i];
    }
    for (int i = 0; i < nl; ++i)
      red.push_back(Rect{sc.TV.get() + (int64_t)i * kk, sc.sp.get() + spo[4 * L[i] + side], ks, ks, ks, rank[i],
                         1.0, 0.0});
    kern::seg_gemm_nn(qt, Ys, N, sc.TU.get(), ks, basis.get(), N, 1.0, 0.0, stream);
    kern::rect_axpby(cp, drc_, stream);
    kern::rect_axpby(red, drc_, stream);
    if (side) rVmax = rmax;
    else rUmax = rmax;
  }
  // omT_i = Omega(I)^T V_i, psT_i = Psi(I)^T U_i
  {
    std::vector<SegTN> to, tp;
    for (int id : L) {
      spo[4 * id + 2] = spool_alloc((int64_t)ks * sn[id].rV);
      spo[4 * id + 3] = spool_alloc((int64_t)ks * sn[id].rU);
    }
    for (int id : L) {
      const TNode& nd = tree.nodes[id];
      to.push_back(SegTN{nd.start, nd.start, spo[4 * id + 2], (int32_t)nd.len, ks, sn[id].rV, 0});
      tp.push_back(SegTN{nd.start, nd.start, spo[4 * id + 3], (int32_t)nd.len, ks, sn[id].rU, 0});
    }
    kern::seg_gemm_tn(to, sc.Om.get(), N, Vl.get(), N, sc.sp.get(), ks, 1.0, 0.0, ws_, stream);
    kern::seg_gemm_tn(tp, sc.Ps.get(), N, Ul.get(), N, sc.sp.get(), ks, 1.0, 0.0, ws_, stream);
  }
  // ---- full bases of one depth: leaf bases, or diag(F_a, F_b) T_p from the
  // child depth's array Fc (T = R for side 0, W for side 1)
  auto make_full = [&](const std::vector<int>& ids, int side, const double* Fc, DevBuf<double>& F) {
    int r = 1, ld = 1;
    for (int id : ids) {
      r = std::max(r, side ? sn[id].rV : sn[id].rU);
      if (!tree.nodes[id].leaf) ld = side ? sn[id].ldW : sn[id].ldR;
    }
    F.reserve((size_t)N * r);
    std::vector<SegNN> cp, mm;
    for (int id : ids) {
      const TNode& nd = tree.nodes[id];
      const int rp = side ? sn[id].rV : sn[id].rU;
      if (nd.leaf) {
        cp.push_back(SegNN{nd.start, 0, nd.start, (int32_t)nd.len, 0, rp, 0});
      } else {
        const TNode& a = tree.nodes[nd.left];
        const TNode& b = tree.nodes[nd.right];
        const int ra = side ? sn[nd.left].rV : sn[nd.left].rU;
        const int rb = side ? sn[nd.right].rV : sn[nd.right].rU;
        const int64_t o = side ? sn[id].oW : sn[id].oR;
        mm.push_back(SegNN{a.start, o, a.start, (int32_t)a.len, ra, rp, 0});
        mm.push_back(SegNN{b.start, o + ra, b.start, (int32_t)b.len, rb, rp, 0});
      }
    }
    kern::copy_segs(cp, dnn_, side ? Vl.get() : Ul.get(), N, F.get(), N, stream);
    kern::seg_gemm_nn(mm, Fc, N, pool.get(), ld, F.get(), N, 1.0, 0.0, stream);
  };
  DevBuf<double>*UFc = &sc.UF[0], *VFc = &sc.VF[0], *UFn = &sc.UF[1], *VFn = &sc.VF[1];
  make_full(tree.by_depth[dmax], 0, nullptr, *UFc);
  make_full(tree.by_depth[dmax], 1, nullptr, *VFc);
  for (int d = dmax - 1; d >= 0; --d) {
    std::vector<int> inner;
    for (int id : tree.by_depth[d])
      if (!tree.nodes[id].leaf) inner.push_back(id);
    const int ni = (int)inner.size();
    // ---- couplings of the children: P(I_a) = A(I_a, I_b) V_b, B_ab = U_a^T P(I_a)
    int ldb = 1, kv = 1;
    for (int id : inner) {
      const TNode& nd = tree.nodes[id];
      ldb = std::max(ldb, std::max(sn[nd.left].rU, sn[nd.right].rU));
      kv = std::max(kv, std::max(sn[nd.left].rV, sn[nd.right].rV));
    }
    for (int id : inner) {
      const TNode& nd = tree.nodes[id];
      sn[id].ldB = ldb;
      sn[id].oBab = pool_alloc((int64_t)ldb * sn[nd.right].rV);
      sn[id].oBba = pool_alloc((int64_t)ldb * sn[nd.left].rV);
    }
    sc.P.reserve((size_t)N * kv);
    {
      std::vector<SegSk> sk;
      std::vector<SegTN> tn;
      for (int id : inner) {
        const TNode& nd = tree.nodes[id];
        const TNode& a = tree.nodes[nd.left];
        const TNode& b = tree.nodes[nd.right];
        const int ua = sn[nd.left].rU, ub = sn[nd.right].rU, va = sn[nd.left].rV, vb = sn[nd.right].rV;
        sk.push_back(SegSk{a.start, b.start, 0, b.start, a.start, (int32_t)a.len, (int32_t)b.len, 0, vb});
        sk.push_back(SegSk{b.start, a.start, 0, a.start, b.start, (int32_t)b.len, (int32_t)a.len, 0, va});
        tn.push_back(SegTN{a.start, a.start, sn[id].oBab, (int32_t)a.len, ua, vb, 0});
        tn.push_back(SegTN{b.start, b.start, sn[id].oBba, (int32_t)b.len, ub, va, 0});
      }
      kern::sketch(g, false, sk, VFc->get(), N, sc.P.get(), N, kv, ws_, stream);
      kern::seg_gemm_tn(tn, UFc->get(), N, sc.P.get(), N, pool.get(), ldb, 1.0, 0.0, ws_, stream);
    }
    if (d == 0) break;  // the root has no transfer
    // ---- local samples: side 0 M_p = [yh_a - B_ab om_b; yh_b - B_ba om_a],
    // side 1 N_p = [zh_a - B_ba^T ps_b; zh_b - B_ab^T ps_a] (m x s).  Tall
    // ones (m > s) are stored as they are in Mt (rows moff_j .., leading
    // dimension Lt) and factored M = Q R; R^T goes to the entry's Jacobi
    // slot.  Small ones go to the slot as M^T zero-padded to s x s.
    const size_t nmat = 2 * (size_t)ni;
    std::vector<int> mrows(nmat);
    std::vector<int64_t> moff(nmat, -1);
    int64_t Lt = 0;
    int ldT = 1;
    for (int i = 0; i < ni; ++i) {
      const TNode& nd = tree.nodes[inner[i]];
      mrows[i] = sn[nd.left].rU + sn[nd.right].rU;
      mrows[ni + i] = sn[nd.left].rV + sn[nd.right].rV;
    }
    for (size_t j = 0; j < nmat; ++j) {
      ldT = std::max(ldT, mrows[j]);
      if (mrows[j] > ks) {
        moff[j] = Lt;
        Lt += mrows[j];
      }
    }
    sc.Sm.reserve(nmat * kk);
    sc.Rt.reserve(nmat * kk);
    sc.Vr.reserve(nmat * kk);
    sc.Sv.reserve(nmat * ks);
    sc.Wk.reserve(2 * nmat * kk);
    sc.TU.reserve(nmat * kk);
    sc.TV.reserve(nmat * kk);
    sc.Mt.reserve((size_t)std::max<int64_t>(Lt, 1) * ks);
    sc.Tb.reserve((size_t)std::max<int64_t>(Lt, 1) * ks);
    zero_bytes_async(sc.Sm.get(), nmat * kk * sizeof(double), stream);
    double* SP = sc.sp.get();
    double* PL = pool.get();
    double* MT = sc.Mt.get();
    std::vector<Rect> cpT, cpS;  // reduced samples into tall M (transposed copy) / small M^T (copy)
    std::vector<BGemm> up;
    for (int i = 0; i < ni; ++i) {
      const int id = inner[i];
      const TNode& nd = tree.nodes[id];
      const int a = nd.left, b = nd.right;
      const int ua = sn[a].rU, ub = sn[b].rU, va = sn[a].rV, vb = sn[b].rV;
      const int64_t ldB = sn[id].ldB;
      const double* Bab = PL + sn[id].oBab;
      const double* Bba = PL + sn[id].oBba;
      for (int side = 0; side < 2; ++side) {
        const size_t j = side ? ni + i : i;
        const int r0 = side ? va : ua, r1 = side ? vb : ub;
        const double* s0 = SP + spo[4 * a + side];
        const double* s1 = SP + spo[4 * b + side];
        // subtracted products (first block, second block): side 0 B_ab om_b,
        // B_ba om_a; side 1 B_ba^T ps_b, B_ab^T ps_a (om, ps stored transposed)
        const double* f0 = side ? Bba : Bab;
        const double* f1 = side ? Bab : Bba;
        const double* g0 = SP + spo[4 * b + (side ? 3 : 2)];
        const double* g1 = SP + spo[4 * a + (side ? 3 : 2)];
        const int k0 = side ? ub : vb, k1 = side ? ua : va;
        if (mrows[j] > ks) {
          double* m = MT + moff[j];
          cpT.push_back(Rect{s0, m, ks, Lt, ks, r0, 1.0, 0.0});
          cpT.push_back(Rect{s1, m + r0, ks, Lt, ks, r1, 1.0, 0.0});
          // M(rows 0:r0) -= op(f0) g0^T
          up.push_back(BGemm{f0, g0, m, r0, ks, k0, (int32_t)ldB, ks, (int32_t)Lt, side, 1, -1.0, 1.0});
          up.push_back(BGemm{f1, g1, m + r0, r1, ks, k1, (int32_t)ldB, ks, (int32_t)Lt, side, 1, -1.0, 1.0});
        } else {
          double* m = sc.Sm.get() + (int64_t)j * kk;
          cpS.push_back(Rect{s0, m, ks, ks, ks, r0, 1.0, 0.0});
          cpS.push_back(Rect{s1, m + (int64_t)r0 * ks, ks, ks, ks, r1, 1.0, 0.0});
          // M^T(:, 0:r0) -= g0 op(f0)^T
          up.push_back(BGemm{g0, f0, m, ks, r0, k0, ks, (int32_t)ldB, ks, 0, side ? 0 : 1, -1.0, 1.0});
          up.push_back(BGemm{g1, f1, m + (int64_t)r0 * ks, ks, r1, k1, ks, (int32_t)ldB, ks, 0, side ? 0 : 1, -1.0,
                             1.0});
        }
      }
    }
    kern::rect_transpose(cpT, drc_, stream);
    kern::rect_axpby(cpS, drc_, stream);
    kern::bgemm(up, dbg_, stream);
    {
      std::vector<OrthSeg> tall;
      std::vector<Rect> rt;
      for (size_t j = 0; j < nmat; ++j)
        if (mrows[j] > ks) {
          rt.push_back(Rect{sc.Rt.get() + (int64_t)tall.size() * kk, sc.Sm.get() + (int64_t)j * kk, ks, ks, ks, ks,
                            1.0, 0.0});
          tall.push_back(OrthSeg{moff[j], mrows[j], ks});
        }
      if (!tall.empty()) orth_.run(tall, MT, sc.Tb.get(), Lt, ks, sc.Rt.get(), stream);
      kern::rect_transpose(rt, drc_, stream);
    }
    std::vector<SmallMat> sm(nmat);
    for (size_t j = 0; j < nmat; ++j) sm[j] = SmallMat{(int64_t)j * kk, ks, ks};
    kern::jacobi_svd(sm, dsm_, sc.Sm.get(), sc.Vr.get(), sc.Sv.get(), ks, sc.Wk.get(), stream);
    const auto S = download(sc.Sv.get(), nmat * ks, stream);
    std::vector<int> rk(nmat);
    int rm = 1;
    for (size_t j = 0; j < nmat; ++j) {
      rk[j] = pick_rank(&S[j * ks], std::min(mrows[j], ks));
      rm = std::max(rm, rk[j]);
    }
    DevBuf<int> drk;
    drk.upload(rk, stream);
    // TU = V(:, :r), TV = U(:, :r) S
    kern::trunc_prep(sm, dsm_, sc.Sm.get(), sc.Vr.get(), sc.Sv.get(), ks, drk.get(), rm, sc.TU.get(), sc.TV.get(),
                     1, stream);
    // transfers (R_p, W_p: Q V(:, :r) for tall samples, V(:m, :r) for small
    // ones) and reduced samples yhT_p, zhT_p = TV(:, :r)
    for (int i = 0; i < ni; ++i) {
      const int id = inner[i];
      sn[id].rU = rk[i];
      sn[id].rV = rk[ni + i];
      sn[id].ldR = ldT;
      sn[id].ldW = ldT;
      sn[id].oR = pool_alloc((int64_t)ldT * rk[i]);
      sn[id].oW = pool_alloc((int64_t)ldT * rk[ni + i]);
      for (int q = 0; q < 4; ++q) spo[4 * id + q] = spool_alloc((int64_t)ks * ((q == 0 || q == 3) ? rk[i] : rk[ni + i]));
    }
    SP = sc.sp.get();
    PL = pool.get();
    std::vector<Rect> st;
    std::vector<SegNN> qv;
    for (int i = 0; i < ni; ++i) {
      const int id = inner[i];
      for (int side = 0; side < 2; ++side) {
        const size_t j = side ? ni + i : i;
        const int64_t ot = side ? sn[id].oW : sn[id].oR;
        if (mrows[j] > ks)
          qv.push_back(SegNN{moff[j], (int64_t)j * kk, ot, mrows[j], ks, rk[j], 0});
        else
          st.push_back(Rect{sc.TU.get() + (int64_t)j * kk, PL + ot, ks, ldT, mrows[j], rk[j], 1.0, 0.0});
        st.push_back(Rect{sc.TV.get() + (int64_t)j * kk, SP + spo[4 * id + side], ks, ks, ks, rk[j], 1.0, 0.0});
      }
    }
    kern::seg_gemm_nn(qv, (const double*)MT, Lt, (const double*)sc.TU.get(), ks, PL, (int64_t)ldT, 1.0, 0.0,
                      stream);
    kern::rect_axpby(st, drc_, stream);
    // omT_p = [omT_a omT_b] W_p, psT_p = [psT_a psT_b] R_p (two launches: one C per item each)
    std::vector<BGemm> r1, r2;
    for (int i = 0; i < ni; ++i) {
      const int id = inner[i];
      const TNode& nd = tree.nodes[id];
      const int a = nd.left, b = nd.right;
      const int ua = sn[a].rU, ub = sn[b].rU, va = sn[a].rV, vb = sn[b].rV;
      r1.push_back(BGemm{SP + spo[4 * a + 2], PL + sn[id].oW, SP + spo[4 * id + 2], ks, sn[id].rV, va, ks, ldT, ks, 0,
                         0, 1.0, 0.0});
      r1.push_back(BGemm{SP + spo[4 * a + 3], PL + sn[id].oR, SP + spo[4 * id + 3], ks, sn[id].rU, ua, ks, ldT, ks, 0,
                         0, 1.0, 0.0});
      r2.push_back(BGemm{SP + spo[4 * b + 2], PL + sn[id].oW + va, SP + spo[4 * id + 2], ks, sn[id].rV, vb, ks, ldT,
                         ks, 0, 0, 1.0, 1.0});
      r2.push_back(BGemm{SP + spo[4 * b + 3], PL + sn[id].oR + ua, SP + spo[4 * id + 3], ks, sn[id].rU, ub, ks, ldT,
                         ks, 0, 0, 1.0, 1.0});
    }
    kern::bgemm(r1, dbg_, stream);
    kern::bgemm(r2, dbg_, stream);
    // ---- full bases of this depth
    make_full(tree.by_depth[d], 0, UFc->get(), *UFn);
    make_full(tree.by_depth[d], 1, VFc->get(), *VFn);
    std::swap(UFc, UFn);
    std::swap(VFc, VFn);
  }
  HIP_CHECK(hipStreamSynchronize(stream));
}

void HssEngine::build(const GenDesc& g, int64_t n, int leaf_size, int mode_, const double* host_omega,
                      int64_t omega_len) {
  HMAT_REQUIRE(n > 0, HMAT_ERR_ARG, "N must be positive");
  HMAT_REQUIRE(leaf_size > 0, HMAT_ERR_ARG, "leaf_size must be positive");
  HMAT_REQUIRE(opt.precision == HMAT_PREC_FP64, HMAT_ERR_ARG, "HSS runs in FP64 only");
  HMAT_REQUIRE(mode_ == kReference || mode_ == kNested || mode_ == kSampled, HMAT_ERR_ARG,
               "HSS build mode must be 0 (reference), 1 (nested) or 2 (sampled)");
  if (mode_ == kSampled) {
    const int ks = sampled_width(opt, n);
    HMAT_REQUIRE(ks <= 256, HMAT_ERR_ARG, "sampled HSS: rank + oversample above 256 is not supported");
    HMAT_REQUIRE(!host_omega || omega_len == sampled_omega_length(opt, n), HMAT_ERR_ARG,
                 "omega_len " + std::to_string(omega_len) + " != expected " +
                     std::to_string(sampled_omega_length(opt, n)));
  }
  built = factored = false;
  N = n;
  leaf = leaf_size;
  mode = mode_;
  nested = mode != kReference;
  tree.build(N, leaf);
  sn.assign(tree.nodes.size(), SNode{});
  pool_used = 0;
  const int64_t lm = tree.lmax;
  const int nl = (int)tree.leaves.size();
  // Omega offsets per leaf side in the serial draw order
  std::vector<int64_t> omega_off(2 * tree.nodes.size(), -1);
  if (host_omega && tree.max_depth > 0 && mode != kSampled) {
    std::vector<int> order = tree.leaves;
    std::sort(order.begin(), order.end(),
              [&](int a, int b) { return tree.nodes[a].ref_id > tree.nodes[b].ref_id; });
    int64_t pos = 0;
    for (int id : order) {
      const int64_t m = tree.nodes[id].len;
      for (int side = 0; side < 2; ++side)
        if (uses_rsvd(m, N, opt)) {
          omega_off[2 * id + side] = pos;
          pos += N * sketch_width(m, N, opt);
        }
    }
    HMAT_REQUIRE(pos == omega_len, HMAT_ERR_ARG,
                 "omega_len " + std::to_string(omega_len) + " != expected " + std::to_string(pos));
  }
  if (tree.max_depth > 0 && mode == kSampled) {
    build_sampled(g, host_omega, omega_len);
  } else if (tree.max_depth > 0) {
    compress_leaves(g, false, host_omega, omega_off);
    compress_leaves(g, true, host_omega, omega_off);
    build_internal();
  }
  // leaf diagonal blocks D (hss_routines_omp.h:75)
  D.resize((size_t)nl * lm * lm);
  D.zero(stream);
  std::vector<BlockDesc> bd;
  for (int i = 0; i < nl; ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    bd.push_back({nd.start, nd.start, (int64_t)i * lm * lm, (int32_t)nd.len, (int32_t)nd.len, (int32_t)lm, 0});
  }
  kern::gen_blocks(g, bd, dbd_, D.get(), stream);
  pool_built = pool_used;
  set_ldv();
  HIP_CHECK(hipStreamSynchronize(stream));
  mv_.ok = false;
  built = true;
}

// Leading dimension of the per-node vectors of the sweeps: wide enough for a
// node's [rU; rV] and for both children's stacked ranks.
void HssEngine::set_ldv() {
  int lv = 1;
  for (auto& x : sn) lv = std::max(lv, std::max(std::max(x.rU, x.rV), x.rU + x.rV));
  for (size_t id = 0; id < tree.nodes.size(); ++id)
    if (!tree.nodes[id].leaf) {
      const TNode& nd = tree.nodes[id];
      lv = std::max(lv, sn[nd.left].rU + sn[nd.right].rU + sn[nd.left].rV + sn[nd.right].rV);
    }
  ldv = lv;
}

double HssEngine::matvec_bytes(int nrhs) const {
  double leaf2 = 0.0, basis = 0.0;
  for (int id : tree.leaves) {
    const double len = (double)tree.nodes[id].len;
    leaf2 += len * len;
    basis += len * (double)(sn[id].rU + sn[id].rV);
  }
  return 8.0 * (leaf2 + basis + (double)pool_built) + 16.0 * (double)N * nrhs;
}

double HssEngine::solve_bytes(int nrhs) const {
  double leaf2 = 0.0, basis = 0.0;
  for (int id : tree.leaves) {
    const double len = (double)tree.nodes[id].len;
    leaf2 += len * len;
    basis += len * (double)(sn[id].rU + sn[id].rV);
  }
  return 8.0 * (2.0 * leaf2 + basis + (double)pool_used) + 16.0 * (double)N * nrhs;
}

// ------------------------------------------------------------------ matvec
// g_i = V_i^T x_i; g_p = W_p^T [g_a; g_b]; f_a = R_p(a) f_p + B_ab g_b;
// y_i = D_i x_i + U_i f_i   (hss_routines_omp.h:154-230)
void HssEngine::matvec(const double* dX, int64_t ldx, int nrhs, double* dY, int64_t ldy) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "matvec before build");
  HMAT_REQUIRE(nrhs > 0 && ldx >= N && ldy >= N, HMAT_ERR_ARG, "bad matvec dimensions");
  ProfScope prof("hss_matvec", stream, matvec_bytes(nrhs));
  if (plan_matvec()) {
    if (kern::hss_mv_launch(mv_, D.get(), Ul.get(), Vl.get(), pool.get(), N, dX, ldx, nrhs, dY, ldy, stream)) {
      prof.finish();
      if (prof_enabled()) prof_add("hss_matvec_one_launch", 0.0, 1.0);
      kern::hss_mv_finish(mv_, stream);
      return;
    }
    mv_.ok = false;  // refused by the runtime: the multi-launch path
  }
  leaf_apply(D.get(), dX, ldx, nrhs, dY, ldy);
  if (tree.max_depth == 0) {
    prof.finish();
    HIP_CHECK(hipStreamSynchronize(stream));
    return;
  }
  const int nn = (int)tree.nodes.size();
  const int64_t vs = (int64_t)ldv * nrhs;
  va_.reserve((size_t)nn * vs);  // g
  vb_.reserve((size_t)nn * vs);  // f
  double* G = va_.get();
  double* F = vb_.get();
  auto gv = [&](int id) { return G + (int64_t)id * vs; };
  auto fv = [&](int id) { return F + (int64_t)id * vs; };
  {
    std::vector<SegTN> tn;
    for (int id : tree.leaves) {
      const TNode& nd = tree.nodes[id];
      tn.push_back(SegTN{nd.start, nd.start, (int64_t)id * vs, (int32_t)nd.len, sn[id].rV, nrhs, 0});
    }
    kern::seg_gemm_tn(tn, Vl.get(), N, dX, ldx, G, ldv, 1.0, 0.0, ws_, stream);
  }
  const double* P = pool.get();
  SweepProgram sw;
  // up-sweep g_p = W_p^T [g_a; g_b], one task per node and phase per depth
  for (int d = tree.max_depth - 1; d >= 1; --d) {
    sw.begin_phase();
    for (int id : tree.by_depth[d]) {
      const TNode& nd = tree.nodes[id];
      if (nd.leaf) continue;
      const int va = sn[nd.left].rV, vb = sn[nd.right].rV;
      sw.begin_task();
      sw.op(BGemm{P + sn[id].oW, gv(nd.left), gv(id), sn[id].rV, nrhs, va, sn[id].ldW, ldv, ldv, 1, 0, 1.0, 0.0});
      sw.op(BGemm{P + sn[id].oW + va, gv(nd.right), gv(id), sn[id].rV, nrhs, vb, sn[id].ldW, ldv, ldv, 1, 0, 1.0, 1.0});
      sw.end_task();
    }
  }
  // down-sweep f_a = B_ab g_b + R_p(a) f_p, f_b = B_ba g_a + R_p(b) f_p
  for (int d = 0; d < tree.max_depth; ++d) {
    sw.begin_phase();
    for (int id : tree.by_depth[d]) {
      const TNode& nd = tree.nodes[id];
      if (nd.leaf) continue;
      const int ua = sn[nd.left].rU, ub = sn[nd.right].rU;
      sw.begin_task();
      sw.op(BGemm{P + sn[id].oBab, gv(nd.right), fv(nd.left), ua, nrhs, sn[nd.right].rV, sn[id].ldB, ldv, ldv, 0, 0, 1.0, 0.0});
      sw.op(BGemm{P + sn[id].oBba, gv(nd.left), fv(nd.right), ub, nrhs, sn[nd.left].rV, sn[id].ldB, ldv, ldv, 0, 0, 1.0, 0.0});
      if (d > 0) {
        sw.op(BGemm{P + sn[id].oR, fv(id), fv(nd.left), ua, nrhs, sn[id].rU, sn[id].ldR, ldv, ldv, 0, 0, 1.0, 1.0});
        sw.op(BGemm{P + sn[id].oR + ua, fv(id), fv(nd.right), ub, nrhs, sn[id].rU, sn[id].ldR, ldv, ldv, 0, 0, 1.0, 1.0});
      }
      sw.end_task();
    }
  }
  sw.run(stream);
  {
    std::vector<SegNN> nn2;
    for (int id : tree.leaves) {
      const TNode& nd = tree.nodes[id];
      nn2.push_back(SegNN{nd.start, (int64_t)id * vs, nd.start, (int32_t)nd.len, sn[id].rU, nrhs, 0});
    }
    kern::seg_gemm_nn(nn2, Ul.get(), N, F, ldv, dY, ldy, 1.0, 1.0, stream);
  }
  prof.finish();
  HIP_CHECK(hipStreamSynchronize(stream));
}

// One-launch matvec descriptors: the leaves, then the sweep's node tasks
// in dependency order (up-sweep deepest depth first, down-sweep from the
// root), as matvec() runs them depth by depth.
bool HssEngine::plan_matvec() {
  if (mv_.ok) return true;
  if (tree.max_depth == 0 || tree.lmax > 1024) return false;
  int rmax = 1;
  for (const SNode& q : sn) rmax = std::max({rmax, q.rU, q.rV});
  if (rmax > 64) return false;
  std::vector<kern::HmvLeaf> lf;
  for (int id : tree.leaves) {
    const TNode& nd = tree.nodes[id];
    lf.push_back(kern::HmvLeaf{nd.start, (int32_t)nd.len, id, sn[id].rU, sn[id].rV});
  }
  std::vector<kern::HmvTask> tk;
  for (int d = tree.max_depth - 1; d >= 1; --d)
    for (int id : tree.by_depth[d]) {
      const TNode& nd = tree.nodes[id];
      if (nd.leaf) continue;
      tk.push_back(kern::HmvTask{0, id, nd.left, nd.right, sn[nd.left].rV, sn[nd.right].rV, sn[id].rV, 0, 0,
                                 sn[id].ldW, 0, 0, sn[id].oW, 0, 0});
    }
  for (int d = 0; d < tree.max_depth; ++d)
    for (int id : tree.by_depth[d]) {
      const TNode& nd = tree.nodes[id];
      if (nd.leaf) continue;
      const int rp = d > 0 ? sn[id].rU : 0;
      tk.push_back(kern::HmvTask{1, id, nd.left, nd.right, sn[nd.left].rU, sn[nd.right].rU, rp, sn[nd.left].rV,
                                 sn[nd.right].rV, sn[id].ldB, sn[id].ldR, 0, sn[id].oBab, sn[id].oBba,
                                 d > 0 ? sn[id].oR : 0});
    }
  mv_.nleaf = (int)lf.size();
  mv_.ntasks = (int)tk.size();
  mv_.nn = (int)tree.nodes.size();
  mv_.ldv = rmax;
  mv_.lm = tree.lmax;
  mv_.leaves.upload(lf, stream);
  mv_.tasks.upload(tk, stream);
  mv_.vec.reserve((size_t)2 * mv_.nn * mv_.ldv * 16 * sizeof(double));
  mv_.sync.reserve(((size_t)2 * mv_.nn + 2) * sizeof(int));
  mv_.ok = true;
  for (int l = 0; l < 5; ++l) mv_.grid[l] = kern::hss_mv_grid(mv_, l);
  if (mv_.grid[0] <= 0) {
    mv_.ok = false;
    return false;
  }
  return true;
}

// -------------------------------------------------------------- factor/solve
// Push-through Woodbury on nested bases (hss_routines_omp.h:306-343):
//   leaf:  Q_i = D_i^{-1} U_i,  K_i = V_i^T Q_i
//   node:  M = I + K C, C = [[0, B_ab], [B_ba, 0]], K = diag(K_a, K_b)
//          G = (I - C M^{-1} K) R,   K_p = W^T K G   (Q_p = diag(Q_a, Q_b) G)
void HssEngine::factor() {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "factor before build");
  factored = false;
  apl_.ok = false;
  ulvp_.ok = false;
  if (solver == kUlv) {
    ulv_factor();
    HIP_CHECK(hipStreamSynchronize(stream));
    factored = true;
    return;
  }
  pool_used = pool_built;  // a refactorisation reuses the factor's part of the pool
  const int64_t lm = tree.lmax;
  const int nl = (int)tree.leaves.size();
  Dinv.resize((size_t)nl * lm * lm);
  HIP_CHECK(hipMemcpyAsync(Dinv.get(), D.get(), (size_t)nl * lm * lm * sizeof(double), hipMemcpyDeviceToDevice, stream));
  std::vector<SmallMat> sm(nl);
  for (int i = 0; i < nl; ++i)
    sm[i] = SmallMat{(int64_t)i * lm * lm, (int32_t)tree.nodes[tree.leaves[i]].len, (int32_t)lm};
  info_.reserve(std::max(nl, 1));
  kern::gj_inverse(sm, dsm_, Dinv.get(), info_.get(), stream);
  auto info = download(info_.get(), nl, stream);
  for (int i = 0; i < nl; ++i)
    if (info[i] != 0) throw Error(HMAT_ERR_SINGULAR, "singular HSS leaf block " + std::to_string(i));
  if (tree.max_depth == 0) {
    factored = true;
    return;
  }
  Ql.resize((size_t)N * std::max(rUmax, 1));
  lres_.reserve((size_t)N * std::max(rUmax, 1));
  leaf_solve(Ul.get(), N, rUmax, Ql.get(), N, lres_.get());
  // K_i = V_i^T Q_i, stored with leading dimension ldv
  for (int id : tree.leaves) sn[id].oK = pool_alloc((int64_t)ldv * sn[id].rU);
  {
    std::vector<SegTN> tn;
    for (int id : tree.leaves) {
      const TNode& nd = tree.nodes[id];
      tn.push_back(SegTN{nd.start, nd.start, sn[id].oK, (int32_t)nd.len, sn[id].rV, sn[id].rU, 0});
    }
    kern::seg_gemm_tn(tn, Vl.get(), N, Ql.get(), N, pool.get(), ldv, 1.0, 0.0, ws_, stream);
  }
  for (int d = tree.max_depth - 1; d >= 0; --d) {
    std::vector<int> inner;
    for (int id : tree.by_depth[d])
      if (!tree.nodes[id].leaf) inner.push_back(id);
    // allocate M, G, K and two scratch blocks per node
    std::vector<int64_t> oX(tree.nodes.size()), oXr(tree.nodes.size()), oC(tree.nodes.size()),
        oKG(tree.nodes.size());
    for (int id : inner) {
      const TNode& nd = tree.nodes[id];
      const int ms = sn[nd.left].rV + sn[nd.right].rV;
      const int kR = sn[nd.left].rU + sn[nd.right].rU;
      sn[id].ms = ms;
      sn[id].oMinv = pool_alloc((int64_t)ms * ms);
      sn[id].oM = pool_alloc((int64_t)ms * ms);
      if (d > 0) {
        sn[id].oG = pool_alloc((int64_t)kR * sn[id].rU);
        sn[id].oK = pool_alloc((int64_t)ldv * sn[id].rU);
        oX[id] = pool_alloc((int64_t)ms * kR);
        oXr[id] = pool_alloc((int64_t)ms * kR);
        oC[id] = pool_alloc((int64_t)kR * kR);
        oKG[id] = pool_alloc((int64_t)ms * sn[id].rU);
      }
    }
    double* P = pool.get();
    // M = I + [[0, K_a B_ab], [K_b B_ba, 0]]
    std::vector<Rect> zm;
    std::vector<BGemm> mb;
    std::vector<SmallMat> ms;
    for (int id : inner) {
      const TNode& nd = tree.nodes[id];
      const int va = sn[nd.left].rV, vb = sn[nd.right].rV, ua = sn[nd.left].rU, ub = sn[nd.right].rU;
      const int m = sn[id].ms;
      double* M = P + sn[id].oMinv;
      zm.push_back(Rect{nullptr, M, 0, m, m, m, 1.0, 0.0});
      mb.push_back(BGemm{P + sn[nd.left].oK, P + sn[id].oBab, M + (int64_t)va * m, va, vb, ua, ldv, sn[id].ldB, m, 0, 0, 1.0, 0.0});
      mb.push_back(BGemm{P + sn[nd.right].oK, P + sn[id].oBba, M + va, vb, va, ub, ldv, sn[id].ldB, m, 0, 0, 1.0, 0.0});
      ms.push_back(SmallMat{sn[id].oMinv, m, m});
    }
    kern::rect_axpby(zm, drc_, stream);
    kern::rect_add_identity(zm, drc_, stream);
    kern::bgemm(mb, dbg_, stream);
    // keep M for the refinement step of every solve with it
    std::vector<Rect> keepM;
    for (int id : inner) {
      const int m = sn[id].ms;
      keepM.push_back(Rect{P + sn[id].oMinv, P + sn[id].oM, m, m, m, m, 1.0, 0.0});
    }
    kern::rect_axpby(keepM, drc_, stream);
    info_.reserve(std::max<size_t>(ms.size(), 1));
    kern::gj_inverse(ms, dsm_, P, info_.get(), stream);
    auto inf = download(info_.get(), ms.size(), stream);
    for (size_t i = 0; i < ms.size(); ++i)
      if (inf[i] != 0) throw Error(HMAT_ERR_SINGULAR, "singular HSS capacitance at depth " + std::to_string(d));
    if (d == 0) break;
    // X = M^{-1} K (ms x kR), refined once: Xr = K - M X, X += M^{-1} Xr;
    // CX = C X (kR x kR);  G = R - CX R;  K_p = W^T (K G)
    std::vector<BGemm> s1, s1r, s1x, s2, s3, s4, s5;
    std::vector<Rect> cpG, addK;
    for (int id : inner) {
      const TNode& nd = tree.nodes[id];
      const int va = sn[nd.left].rV, vb = sn[nd.right].rV, ua = sn[nd.left].rU, ub = sn[nd.right].rU;
      const int m = sn[id].ms, kR = ua + ub, up = sn[id].rU;
      const double* Mi = P + sn[id].oMinv;
      double* X = P + oX[id];
      double* CX = P + oC[id];
      double* G = P + sn[id].oG;
      double* KG = P + oKG[id];
      s1.push_back(BGemm{Mi, P + sn[nd.left].oK, X, m, ua, va, m, ldv, m, 0, 0, 1.0, 0.0});
      s1.push_back(BGemm{Mi + (int64_t)va * m, P + sn[nd.right].oK, X + (int64_t)ua * m, m, ub, vb, m, ldv, m, 0, 0, 1.0, 0.0});
      double* Xr = P + oXr[id];
      s1r.push_back(BGemm{P + sn[id].oM, X, Xr, m, kR, m, m, m, m, 0, 0, -1.0, 0.0});
      addK.push_back(Rect{P + sn[nd.left].oK, Xr, ldv, m, va, ua, 1.0, 1.0});
      addK.push_back(Rect{P + sn[nd.right].oK, Xr + (int64_t)ua * m + va, ldv, m, vb, ub, 1.0, 1.0});
      s1x.push_back(BGemm{Mi, Xr, X, m, kR, m, m, m, m, 0, 0, 1.0, 1.0});
      s2.push_back(BGemm{P + sn[id].oBab, X + va, CX, ua, kR, vb, sn[id].ldB, m, kR, 0, 0, 1.0, 0.0});
      s2.push_back(BGemm{P + sn[id].oBba, X, CX + ua, ub, kR, va, sn[id].ldB, m, kR, 0, 0, 1.0, 0.0});
      cpG.push_back(Rect{P + sn[id].oR, G, sn[id].ldR, kR, kR, up, 1.0, 0.0});
      s3.push_back(BGemm{CX, P + sn[id].oR, G, kR, up, kR, kR, sn[id].ldR, kR, 0, 0, -1.0, 1.0});
      s4.push_back(BGemm{P + sn[nd.left].oK, G, KG, va, up, ua, ldv, kR, m, 0, 0, 1.0, 0.0});
      s4.push_back(BGemm{P + sn[nd.right].oK, G + ua, KG + va, vb, up, ub, ldv, kR, m, 0, 0, 1.0, 0.0});
      s5.push_back(BGemm{P + sn[id].oW, KG, P + sn[id].oK, sn[id].rV, up, m, sn[id].ldW, m, ldv, 1, 0, 1.0, 0.0});
    }
    kern::bgemm(s1, dbg_, stream);
    kern::bgemm(s1r, dbg_, stream);
    kern::rect_axpby(addK, drc_, stream);
    kern::bgemm(s1x, dbg_, stream);
    kern::bgemm(s2, dbg_, stream);
    kern::rect_axpby(cpG, drc_, stream);
    kern::bgemm(s3, dbg_, stream);
    kern::bgemm(s4, dbg_, stream);
    kern::bgemm(s5, dbg_, stream);
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  factored = true;
}

// x_i = D_i^{-1} b_i, h_i = V_i^T x_i; up: u = M^{-1} [h_a; h_b],
// s = C u, h_p = W^T ([h_a; h_b] - K s); down: tau_children = s + G tau_p;
// leaves: x_i -= Q_i tau_i.
// X = A^{-1} B with FP64 iterative refinement through the HSS matvec when
// refine_rtol > 0: r = B - A X, X += A^{-1} r until max_j ||r_j|| / ||b_j||
// <= refine_rtol, refine_max corrections, or a correction that does not
// reduce the residual (the push-through solve amplifies rounding through
// its capacitance systems; refinement takes the residual to the rounding
// floor of the FP64 product).
void HssEngine::solve(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "solve before build");
  HMAT_REQUIRE(nrhs > 0 && ldb >= N && ldx >= N, HMAT_ERR_ARG, "bad solve dimensions");
  if (!factored) factor();
  apply_inverse(dB, ldb, nrhs, dX, ldx);
  last_iters = 0;
  last_resid = -1.0;
  if (!(refine_rtol > 0.0)) return;
  ir_y_.reserve((size_t)N * nrhs);
  ir_r_.reserve((size_t)N * nrhs);
  ir_n_.reserve((size_t)2 * nrhs);
  ir_p_.reserve((size_t)64 * nrhs);
  kern::col_norms2(dB, ldb, N, nrhs, ir_n_.get(), ir_p_.get(), stream);
  const auto bn = download(ir_n_.get(), nrhs, stream);
  double resid = 0.0, prev = 0.0;
  int it = 0;
  ir_best_.reserve((size_t)N * nrhs);
  for (;; ++it) {
    matvec(dX, ldx, nrhs, ir_y_.get(), N);
    kern::axpby_cols(N, nrhs, 1.0, dB, ldb, -1.0, ir_y_.get(), N, ir_r_.get(), N, stream);
    kern::col_norms2(ir_r_.get(), N, N, nrhs, ir_n_.get(), ir_p_.get(), stream);
    const auto rn = download(ir_n_.get(), nrhs, stream);
    resid = 0.0;
    for (int j = 0; j < nrhs; ++j) {
      const double rj = bn[j] > 0.0 ? std::sqrt(rn[j] / bn[j]) : std::sqrt(rn[j]);
      if (!std::isfinite(rj))  // (std::max would drop a NaN)
        throw Error(HMAT_ERR_NONFINITE, "non-finite residual after " + std::to_string(it) +
                                            " refinement iterations");
      resid = std::max(resid, rj);
    }
    // a correction that did not reduce the residual: back to the previous
    // iterate, the one with the lowest residual
    if (it > 0 && !(resid < prev)) {
      HIP_CHECK(hipMemcpy2DAsync(dX, ldx * sizeof(double), ir_best_.get(), N * sizeof(double),
                                 N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream));
      resid = prev;
      --it;
      break;
    }
    if (resid <= refine_rtol || it >= refine_max) break;
    prev = resid;
    HIP_CHECK(hipMemcpy2DAsync(ir_best_.get(), N * sizeof(double), dX, ldx * sizeof(double),
                               N * sizeof(double), nrhs, hipMemcpyDeviceToDevice, stream));
    apply_inverse(ir_r_.get(), N, nrhs, ir_y_.get(), N);
    kern::axpby_cols(N, nrhs, 1.0, dX, ldx, 1.0, ir_y_.get(), N, dX, ldx, stream);
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  last_iters = it;
  last_resid = resid;
  if (resid > refine_rtol)
    throw Error(HMAT_ERR_NOT_CONVERGED, "refinement stopped at relative residual " + sci(resid) +
                                            " after " + std::to_string(it) + " iterations (target " +
                                            sci(refine_rtol) + ")");
}

// Descriptors of the Woodbury apply-inverse for nrhs columns: the leaf
// blocks' segments, V^T X over the leaves, the sweep program (up-sweep
// then down-sweep, one phase per depth) and the leaf expansion, built and
// uploaded once and kept while the column count and the per-node vectors
// stay the same (rebuilt after every factorisation).
void HssEngine::plan_apply(int nrhs) {
  apl_.ok = false;
  const int64_t lm = tree.lmax;
  std::vector<SegNN> leafs;
  for (size_t i = 0; i < tree.leaves.size(); ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    leafs.push_back(SegNN{(int64_t)i * lm * lm, nd.start, nd.start, (int32_t)nd.len, (int32_t)nd.len, nrhs, 0});
  }
  kern::plan_nn(leafs, apl_.leaf, stream);
  apl_.phases.clear();
  if (tree.max_depth > 0) {
    const int64_t vs = (int64_t)ldv * nrhs;
    double *H = va_.get(), *U = vb_.get(), *Sv = vc_.get(), *T = ve_.get(), *Rm = vf_.get();
    auto at = [&](double* b, int id) { return b + (int64_t)id * vs; };
    std::vector<SegTN> tn;
    std::vector<SegNN> up;
    for (int id : tree.leaves) {
      const TNode& nd = tree.nodes[id];
      tn.push_back(SegTN{nd.start, nd.start, (int64_t)id * vs, (int32_t)nd.len, sn[id].rV, nrhs, 0});
      up.push_back(SegNN{nd.start, (int64_t)id * vs, nd.start, (int32_t)nd.len, sn[id].rU, nrhs, 0});
    }
    kern::plan_tn(tn, apl_.vt, stream);
    kern::plan_nn(up, apl_.expand, stream);
    const double* P = pool.get();
    SweepProgram sw;
    // up-sweep, one task per node: u = M^{-1} [h_a; h_b] refined once against
    // M, s = [B_ab u_b; B_ba u_a], and for non-root nodes h_p = W^T u.  The
    // last is the push-through formula's h_p = W^T ([h_a; h_b] - K s) with
    // K s = K C u = (M - I) u and M u = [h_a; h_b]: three dependent products
    // per node fewer than forming K s.
    for (int d = tree.max_depth - 1; d >= 0; --d) {
      sw.begin_phase();
      for (int id : tree.by_depth[d]) {
        const TNode& nd = tree.nodes[id];
        if (nd.leaf) continue;
        const int va = sn[nd.left].rV, vb = sn[nd.right].rV, ua = sn[nd.left].rU, ub = sn[nd.right].rU;
        const int m = sn[id].ms;
        const double* Mi = P + sn[id].oMinv;
        sw.begin_task();
        sw.copy(at(H, nd.left), ldv, at(Rm, id), ldv, va, nrhs);
        sw.copy(at(H, nd.right), ldv, at(Rm, id) + va, ldv, vb, nrhs);
        sw.op(BGemm{Mi, at(Rm, id), at(U, id), m, nrhs, m, m, ldv, ldv, 0, 0, 1.0, 0.0});
        sw.op(BGemm{P + sn[id].oM, at(U, id), at(Rm, id), m, nrhs, m, m, ldv, ldv, 0, 0, -1.0, 1.0});
        sw.op(BGemm{Mi, at(Rm, id), at(U, id), m, nrhs, m, m, ldv, ldv, 0, 0, 1.0, 1.0});
        sw.op(BGemm{P + sn[id].oBab, at(U, id) + va, at(Sv, id), ua, nrhs, vb, sn[id].ldB, ldv, ldv, 0, 0, 1.0, 0.0});
        sw.op(BGemm{P + sn[id].oBba, at(U, id), at(Sv, id) + ua, ub, nrhs, va, sn[id].ldB, ldv, ldv, 0, 0, 1.0, 0.0});
        if (d > 0)
          sw.op(BGemm{P + sn[id].oW, at(U, id), at(H, id), sn[id].rV, nrhs, m, sn[id].ldW, ldv, ldv, 1, 0, 1.0, 0.0});
        sw.end_task();
      }
    }
    // down-sweep, one task per node: t_a = s_a + G_p(a) t_p, t_b = s_b + G_p(b) t_p
    for (int d = 0; d < tree.max_depth; ++d) {
      sw.begin_phase();
      for (int id : tree.by_depth[d]) {
        const TNode& nd = tree.nodes[id];
        if (nd.leaf) continue;
        const int ua = sn[nd.left].rU, ub = sn[nd.right].rU, kR = ua + ub;
        sw.begin_task();
        sw.copy(at(Sv, id), ldv, at(T, nd.left), ldv, ua, nrhs);
        sw.copy(at(Sv, id) + ua, ldv, at(T, nd.right), ldv, ub, nrhs);
        if (d > 0) {
          sw.op(BGemm{P + sn[id].oG, at(T, id), at(T, nd.left), ua, nrhs, sn[id].rU, kR, ldv, ldv, 0, 0, 1.0, 1.0});
          sw.op(BGemm{P + sn[id].oG + ua, at(T, id), at(T, nd.right), ub, nrhs, sn[id].rU, kR, ldv, ldv, 0, 0, 1.0, 1.0});
        }
        sw.end_task();
      }
    }
    apl_.ops.upload(sw.ops, stream);
    apl_.tasks.upload(sw.tasks, stream);
    apl_.phases = sw.phases;
  }
  apl_.nrhs = nrhs;
  apl_.addrs = {pool.get(), va_.get(), vb_.get(), vc_.get(), ve_.get(), vf_.get()};
  apl_.ok = true;
}

void HssEngine::apply_inverse(const double* dB, int64_t ldb, int nrhs, double* dX, int64_t ldx) {
  if (solver == kUlv) {
    ulv_apply(dB, ldb, nrhs, dX, ldx);
    return;
  }
  ProfScope prof("hss_solve_apply", stream, solve_bytes(nrhs));
  const int64_t lm = tree.lmax;
  const int nn = (int)tree.nodes.size();
  const int64_t vs = (int64_t)ldv * nrhs;
  lres_.reserve((size_t)N * nrhs);
  if (tree.max_depth > 0) {
    va_.reserve((size_t)nn * vs);  // h
    vb_.reserve((size_t)nn * vs);  // u
    vc_.reserve((size_t)nn * vs);  // s
    ve_.reserve((size_t)nn * vs);  // tau
    vf_.reserve((size_t)nn * vs);  // [h_a; h_b], then the residual of the M solve
  }
  const std::vector<const void*> addrs = {pool.get(), va_.get(), vb_.get(), vc_.get(), ve_.get(), vf_.get()};
  if (!apl_.ok || apl_.nrhs != nrhs || apl_.addrs != addrs) plan_apply(nrhs);
  // leaf solves, refined once against the leaf blocks (as leaf_solve)
  kern::launch_nn<double>(apl_.leaf, Dinv.get(), lm, dB, ldb, dX, ldx, 1.0, 0.0, stream);
  kern::convert_cols<double, double>(N, nrhs, dB, ldb, lres_.get(), N, stream);
  kern::launch_nn<double>(apl_.leaf, D.get(), lm, dX, ldx, lres_.get(), N, -1.0, 1.0, stream);
  kern::launch_nn<double>(apl_.leaf, Dinv.get(), lm, lres_.get(), N, dX, ldx, 1.0, 1.0, stream);
  if (tree.max_depth > 0) {
    DevBuf<double>& part = ws_.parts<double>();
    part.reserve((size_t)std::max(apl_.vt.nslots, 1) * 4096);
    kern::launch_tn<double>(apl_.vt, Vl.get(), N, dX, ldx, va_.get(), ldv, 1.0, 0.0, part.get(), stream);
    for (const auto& ph : apl_.phases)
      kern::node_tasks(apl_.tasks.get() + ph.first, ph.second - ph.first, apl_.ops.get(), stream);
    kern::launch_nn<double>(apl_.expand, Ql.get(), N, ve_.get(), ldv, dX, ldx, -1.0, 1.0, stream);
  }
  prof.finish();
  HIP_CHECK(hipStreamSynchronize(stream));
}

// ------------------------------------------------------------------ dense
void HssEngine::to_dense_host(double* A, int64_t lda) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "to_dense before build");
  const int64_t lm = tree.lmax;
  auto d = download(D.get(), D.size(), stream);
  for (size_t i = 0; i < tree.leaves.size(); ++i) {
    const TNode& nd = tree.nodes[tree.leaves[i]];
    for (int64_t c = 0; c < nd.len; ++c)
      for (int64_t r = 0; r < nd.len; ++r) A[(nd.start + r) + (nd.start + c) * lda] = d[i * lm * lm + r + c * lm];
  }
  if (tree.max_depth == 0) return;
  auto P = download(pool.get(), (size_t)pool_used, stream);
  auto ul = download(Ul.get(), Ul.size(), stream);
  auto vl = download(Vl.get(), Vl.size(), stream);
  const int nn = (int)tree.nodes.size();
  // full bases, bottom-up: UF[id] is len x rU (column-major)
  std::vector<std::vector<double>> UF(nn), VF(nn);
  for (int dd = tree.max_depth; dd >= 1; --dd)
    for (int id : tree.by_depth[dd]) {
      const TNode& nd = tree.nodes[id];
      const int ru = sn[id].rU, rv = sn[id].rV;
      UF[id].assign((size_t)nd.len * ru, 0.0);
      VF[id].assign((size_t)nd.len * rv, 0.0);
      if (nd.leaf) {
        for (int c = 0; c < ru; ++c)
          for (int64_t r = 0; r < nd.len; ++r) UF[id][r + c * nd.len] = ul[nd.start + r + (int64_t)c * N];
        for (int c = 0; c < rv; ++c)
          for (int64_t r = 0; r < nd.len; ++r) VF[id][r + c * nd.len] = vl[nd.start + r + (int64_t)c * N];
        continue;
      }
      const int a = nd.left, b = nd.right;
      const int64_t la = tree.nodes[a].len, lb = tree.nodes[b].len;
      for (int side = 0; side < 2; ++side) {
        auto& out = side ? VF[id] : UF[id];
        const auto& fa = side ? VF[a] : UF[a];
        const auto& fb = side ? VF[b] : UF[b];
        const int ra = side ? sn[a].rV : sn[a].rU, rb = side ? sn[b].rV : sn[b].rU;
        const int rp = side ? rv : ru;
        const double* Tm = &P[side ? sn[id].oW : sn[id].oR];
        const int ldt = side ? sn[id].ldW : sn[id].ldR;
        for (int c = 0; c < rp; ++c) {
          for (int64_t r = 0; r < la; ++r) {
            double acc = 0.0;
            for (int t = 0; t < ra; ++t) acc += fa[r + t * la] * Tm[t + (int64_t)c * ldt];
            out[r + c * nd.len] = acc;
          }
          for (int64_t r = 0; r < lb; ++r) {
            double acc = 0.0;
            for (int t = 0; t < rb; ++t) acc += fb[r + t * lb] * Tm[ra + t + (int64_t)c * ldt];
            out[la + r + c * nd.len] = acc;
          }
        }
      }
    }
  for (int id = 0; id < nn; ++id) {
    const TNode& nd = tree.nodes[id];
    if (nd.leaf) continue;
    for (int side = 0; side < 2; ++side) {
      const int rr = side ? nd.right : nd.left, cc = side ? nd.left : nd.right;
      const TNode& rn = tree.nodes[rr];
      const TNode& cn = tree.nodes[cc];
      const double* B = &P[side ? sn[id].oBba : sn[id].oBab];
      const int ldb = sn[id].ldB, ru = sn[rr].rU, rv = sn[cc].rV;
      std::vector<double> UB((size_t)rn.len * rv, 0.0);
      for (int c = 0; c < rv; ++c)
        for (int64_t r = 0; r < rn.len; ++r) {
          double acc = 0.0;
          for (int t = 0; t < ru; ++t) acc += UF[rr][r + t * rn.len] * B[t + (int64_t)c * ldb];
          UB[r + c * rn.len] = acc;
        }
      for (int64_t j = 0; j < cn.len; ++j)
        for (int64_t i = 0; i < rn.len; ++i) {
          double acc = 0.0;
          for (int t = 0; t < rv; ++t) acc += UB[i + t * rn.len] * VF[cc][j + t * cn.len];
          A[(rn.start + i) + (cn.start + j) * lda] = acc;
        }
    }
  }
}

}  // namespace hmat

namespace hmat {

// ------------------------------------------------------------ operator file
// Version 1 (native byte order and struct layout):
//   char magic[8] "HMATHSS"; int32 version
//   hmat_options opt; int64 N; int32 leaf; int32 nested
//   int32 nnodes; int32 rUmax; int32 rVmax
//   per tree node (pre-order ids): int32 rU, rV, ldR, ldW, ldB;
//                                  int64 oR, oW, oBab, oBba
//   int64 npool; double pool[npool]   (transfers R/W and couplings B)
//   double Ul[N * rUmax], Vl[N * rVmax]   (leaf bases, column-major, ld N)
//   double D[nleaf * lmax * lmax]
// The tree is rebuilt from (N, leaf); the factorisation is not stored.
static const char kHssMagic[8] = {'H', 'M', 'A', 'T', 'H', 'S', 'S', 0};

void HssEngine::save(const std::string& path) {
  HMAT_REQUIRE(built, HMAT_ERR_STATE, "save before build");
  std::ofstream f(path, std::ios::binary);
  HMAT_REQUIRE(f.good(), HMAT_ERR_ARG, "cannot open " + path + " for writing");
  auto put = [&](const void* p, size_t n) { f.write(reinterpret_cast<const char*>(p), (std::streamsize)n); };
  const int32_t version = 1, lf = leaf, ne = nested ? 1 : 0, nn = (int32_t)tree.nodes.size();
  put(kHssMagic, sizeof(kHssMagic));
  put(&version, 4);
  put(&opt, sizeof(opt));
  put(&N, 8);
  put(&lf, 4);
  put(&ne, 4);
  put(&nn, 4);
  put(&rUmax, 4);
  put(&rVmax, 4);
  for (const SNode& x : sn) {
    const int32_t iv[5] = {x.rU, x.rV, x.ldR, x.ldW, x.ldB};
    const int64_t ov[4] = {x.oR, x.oW, x.oBab, x.oBba};
    put(iv, sizeof(iv));
    put(ov, sizeof(ov));
  }
  put(&pool_built, 8);
  auto dump = [&](const DevBuf<double>& b, size_t n) {
    if (n == 0) return;
    auto h = download(b.get(), n, stream);
    put(h.data(), 8 * n);
  };
  const bool tiny = tree.max_depth == 0;
  dump(pool, tiny ? 0 : (size_t)pool_built);
  dump(Ul, tiny ? 0 : (size_t)N * rUmax);
  dump(Vl, tiny ? 0 : (size_t)N * rVmax);
  dump(D, (size_t)tree.leaves.size() * tree.lmax * tree.lmax);
  HMAT_REQUIRE(f.good(), HMAT_ERR_ARG, "write to " + path + " failed");
}

void HssEngine::load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  HMAT_REQUIRE(f.good(), HMAT_ERR_ARG, "cannot open " + path);
  auto get = [&](void* p, size_t n) {
    f.read(reinterpret_cast<char*>(p), (std::streamsize)n);
    HMAT_REQUIRE(f.good(), HMAT_ERR_ARG, path + ": truncated operator file");
  };
  char magic[8];
  int32_t version = 0;
  get(magic, 8);
  get(&version, 4);
  HMAT_REQUIRE(std::memcmp(magic, kHssMagic, 8) == 0 && version == 1, HMAT_ERR_ARG,
               path + ": not a version-1 HSS operator file");
  hmat_options o;
  int64_t n = 0;
  int32_t lf = 0, ne = 0, nn = 0, ru = 0, rv = 0;
  get(&o, sizeof(o));
  get(&n, 8);
  get(&lf, 4);
  get(&ne, 4);
  get(&nn, 4);
  get(&ru, 4);
  get(&rv, 4);
  HMAT_REQUIRE(n > 0 && lf > 0 && (ne == 0 || ne == 1) && ru >= 0 && rv >= 0 && o.precision == HMAT_PREC_FP64,
               HMAT_ERR_ARG, path + ": bad header");
  Tree t;
  t.build(n, lf);
  HMAT_REQUIRE(nn == (int32_t)t.nodes.size(), HMAT_ERR_ARG, path + ": tree does not match the header");
  std::vector<SNode> s(nn);
  for (SNode& x : s) {
    int32_t iv[5];
    int64_t ov[4];
    get(iv, sizeof(iv));
    get(ov, sizeof(ov));
    x.rU = iv[0], x.rV = iv[1], x.ldR = iv[2], x.ldW = iv[3], x.ldB = iv[4];
    x.oR = ov[0], x.oW = ov[1], x.oBab = ov[2], x.oBba = ov[3];
  }
  int64_t np = 0;
  get(&np, 8);
  HMAT_REQUIRE(np >= 0, HMAT_ERR_ARG, path + ": bad pool size");
  // every generator must lie inside the pool / the leaf bases
  auto inside = [&](int64_t off, int64_t ld, int64_t rows, int64_t cols) {
    if (rows == 0 || cols == 0) return true;
    return off >= 0 && ld >= rows && off + ld * (cols - 1) + rows <= np;
  };
  for (int id = 0; id < nn; ++id) {
    const TNode& nd = t.nodes[id];
    const SNode& x = s[id];
    bool ok = x.rU >= 0 && x.rV >= 0;
    if (nd.leaf) {
      ok = ok && x.rU <= ru && x.rV <= rv && x.rU <= nd.len && x.rV <= nd.len;
    } else if (t.max_depth > 0) {
      const SNode& a = s[nd.left];
      const SNode& b = s[nd.right];
      ok = ok && x.rU <= a.rU + b.rU && x.rV <= a.rV + b.rV;
      if (id != 0)  // the root's transfers are never applied
        ok = ok && inside(x.oR, x.ldR, a.rU + b.rU, x.rU) && inside(x.oW, x.ldW, a.rV + b.rV, x.rV);
      ok = ok && inside(x.oBab, x.ldB, a.rU, b.rV) && inside(x.oBba, x.ldB, b.rU, a.rV);
    }
    HMAT_REQUIRE(ok, HMAT_ERR_ARG, path + ": bad generator layout at node " + std::to_string(id));
  }
  const bool tiny = t.max_depth == 0;
  HMAT_REQUIRE(!tiny || np == 0, HMAT_ERR_ARG, path + ": bad pool size");
  std::vector<double> hp(tiny ? 0 : (size_t)np), hu(tiny ? 0 : (size_t)n * ru), hv(tiny ? 0 : (size_t)n * rv),
      hd((size_t)t.leaves.size() * t.lmax * t.lmax);
  get(hp.data(), 8 * hp.size());
  get(hu.data(), 8 * hu.size());
  get(hv.data(), 8 * hv.size());
  get(hd.data(), 8 * hd.size());
  // the file is whole: replace the handle's operator
  built = factored = false;
  opt = o;
  N = n;
  leaf = lf;
  nested = ne != 0;
  tree = std::move(t);
  sn = std::move(s);
  rUmax = ru;
  rVmax = rv;
  pool_used = pool_built = np;
  auto put_dev = [&](DevBuf<double>& b, const std::vector<double>& h) {
    b.resize(std::max<size_t>(h.size(), 1));
    if (!h.empty()) b.upload(h, stream);
  };
  put_dev(pool, hp);
  put_dev(Ul, hu);
  put_dev(Vl, hv);
  put_dev(D, hd);
  HIP_CHECK(hipStreamSynchronize(stream));
  set_ldv();
  mv_.ok = false;
  built = true;
}

}  // namespace hmat
