"""Build + factor + solve of a batch recorded into one HIP graph
(hmat_hodlr_batch_solve_device with use_graph = 1).

The first call for a batch shape runs the fixed schedule eagerly (choosing
the number of CholeskyQR passes of every orthogonalisation) and records it;
later calls with the same shape replay the graph.  Checked here: on the
recorded values the first call and every replay agree bit for bit with the
eager run of a fresh handle; on new values (Gaussian widths written into the
points in place, a new diagonal shift) the replays agree with it to the
compression error, with eager device work and host synchronisations between
the replays; every
system's residual against its compressed operator stays within ten times
the refinement target (which bounds each column of the whole batch); a replay whose recorded refinement count falls short of a tighter
target is refined further eagerly; a new shape, a transposed solve or a
solve with more columns (buffers that move) records anew; the graph's
operator is the adaptive eager build's to the compression error."""
import numpy as np
import pytest
import torch

import hmat_amd

pytestmark = pytest.mark.gpu


def _opts():
    return hmat_amd.make_options(0.0, rank=24, fixed_rank=True, oversample=10, power_iters=1,
                                 exact_below=0, omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=99)


def _batch(nsys, nlo, nhi, seed, cols):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(nlo), np.log(nhi), nsys)).astype(np.int64)
    pts = [np.sort(rng.random(int(n))) for n in sizes]
    n = int(sizes.sum())
    b = torch.tensor(rng.standard_normal((cols, n)), device="cuda")  # column-major n x cols
    return [int(v) for v in sizes], pts, n, b


def _points(pts, sigmas):
    return np.concatenate([p / s for p, s in zip(pts, sigmas)])


class Batch:
    """A handle solving the batch through the graph entry point."""

    def __init__(self, dp, lam, rtol=1e-11):
        self.h = hmat_amd.HODLRMatrixOMP(opts=_opts())
        # rtol None: no refinement (any residual below 1 accepted, no steps)
        self.h.set_refinement(rtol, 10) if rtol is not None else self.h.set_refinement(1.0, 0)
        self.dp = dp
        self.lam = lam

    def solve(self, sizes, b, graph=True):
        n, cols = b.shape[1], b.shape[0]
        x = torch.empty_like(b)
        k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_GAUSSIAN, self.dp.data_ptr(), p0=1.0, p1=self.lam)
        torch.cuda.synchronize()
        path = self.h.batch_solve_device(k, sizes, 64, b.data_ptr(), n, cols, x.data_ptr(), n, graph=graph)
        torch.cuda.synchronize()
        if graph and path in ("eager", "redone"):
            print(f"\nbatch call: {path}: {self.h.batch_note()}", flush=True)
        return path, x


def _eager_fixed(values, lam, sizes, b, rtol=1e-11):
    """The fixed schedule run eagerly: the first call of a fresh handle."""
    dp = torch.tensor(values, device="cuda")
    path, x = Batch(dp, lam, rtol).solve(sizes, b)
    assert path == "recorded"
    return x


def _residuals(h, sizes, b, x):
    n, cols = b.shape[1], b.shape[0]
    y = torch.empty_like(b)
    h.multiply_device(x.data_ptr(), n, cols, y.data_ptr(), n)
    torch.cuda.synchronize()
    r = (b - y).cpu().numpy()
    bh = b.cpu().numpy()
    offs = np.concatenate([[0], np.cumsum(sizes)])
    worst = 0.0
    for s in range(len(sizes)):
        sl = slice(offs[s], offs[s + 1])
        for j in range(cols):
            worst = max(worst, np.linalg.norm(r[j, sl]) / np.linalg.norm(bh[j, sl]))
    return worst


def test_replays_match_eager_fixed_schedule():
    """Without refinement the recorded schedule ends with the apply alone.
    A replay on the recorded values equals a fresh handle's eager run bit
    for bit; on new values it keeps the recorded number of CholeskyQR passes
    per call, so it agrees with a fresh handle's run (whose pass counts are
    chosen for the new values) to the compression error, and its residual
    is that of an unrefined FP64 solve."""
    sizes, pts, n, b = _batch(12, 400, 3000, 7, 8)
    rng = np.random.default_rng(8)
    sig = rng.uniform(0.05, 0.3, len(sizes))
    dp = torch.tensor(_points(pts, sig), device="cuda")
    bt = Batch(dp, 1e-2, rtol=None)
    path, x0 = bt.solve(sizes, b)
    assert path == "recorded"
    assert torch.equal(x0, _eager_fixed(_points(pts, sig), 1e-2, sizes, b, rtol=None))
    # the same values again, then new widths and new shifts, with eager work
    # and synchronisations in between
    cases = [(sig, 1e-2), (rng.uniform(0.05, 0.3, len(sizes)), 1e-2),
             (rng.uniform(0.05, 0.3, len(sizes)), 3e-2), (sig, 1e-2)]
    for i, (s_i, lam) in enumerate(cases):
        dp.copy_(torch.tensor(_points(pts, s_i), device="cuda"))
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        bt.lam = lam
        path, x = bt.solve(sizes, b)
        ref = _eager_fixed(_points(pts, s_i), lam, sizes, b, rtol=None)
        worst = _residuals(bt.h, sizes, b, x)
        print(f"\nreplay {i}: lam {lam:g}: path {path}, max |x - eager| {float((x - ref).abs().max()):.1e}, "
              f"worst system residual {worst:.1e}", flush=True)
        assert path == "replayed"
        if s_i is sig:
            assert torch.equal(x, ref)
        else:
            assert float(torch.linalg.norm(x - ref) / torch.linalg.norm(ref)) < 1e-7
        assert worst <= 1e-9
        del junk


def test_refined_replays_meet_target():
    """With a 1e-11 target (per column of the whole batch) each replay, or
    its eager finish, leaves every system and column below 1e-10, and agrees
    with the fresh handle's solution to the target times the conditioning."""
    sizes, pts, n, b = _batch(12, 400, 3000, 9, 8)
    rng = np.random.default_rng(10)
    sig = rng.uniform(0.05, 0.3, len(sizes))
    dp = torch.tensor(_points(pts, sig), device="cuda")
    bt = Batch(dp, 1e-2)
    path, x0 = bt.solve(sizes, b)
    assert path == "recorded"
    assert bt.h.last_refinement()[1] <= 1e-11
    assert _residuals(bt.h, sizes, b, x0) <= 1e-10
    for i, lam in enumerate((1e-2, 2e-2, 1e-2)):
        s_i = rng.uniform(0.05, 0.3, len(sizes))
        dp.copy_(torch.tensor(_points(pts, s_i), device="cuda"))
        torch.cuda.synchronize()
        bt.lam = lam
        path, x = bt.solve(sizes, b)
        worst = _residuals(bt.h, sizes, b, x)
        ref = _eager_fixed(_points(pts, s_i), lam, sizes, b)
        err = float(torch.linalg.norm(x - ref) / torch.linalg.norm(ref))
        print(f"\nrefined replay {i}: {path}, {bt.h.last_refinement()}, worst residual {worst:.1e}, "
              f"vs eager {err:.1e}", flush=True)
        assert path in ("replayed", "redone")
        assert bt.h.last_refinement()[1] <= 1e-11
        assert worst <= 1e-10
        assert err < 1e-8


def test_graph_operator_matches_adaptive_build():
    """The fixed schedule (shifted CholeskyQR3 bases) and the adaptive one
    compress each system's kernel matrix equally well: per system, the
    recorded build's error against the dense matrix is within twice the
    adaptive build's."""
    from oracle import kernels
    sizes, pts, n, b = _batch(6, 500, 2000, 11, 4)
    sig = np.random.default_rng(12).uniform(0.05, 0.3, len(sizes))
    dp = torch.tensor(_points(pts, sig), device="cuda")
    bg = Batch(dp, 1e-2)
    be = Batch(dp, 1e-2)
    assert bg.solve(sizes, b)[0] == "recorded"
    be.solve(sizes, b, graph=False)
    ag, ae = bg.h.to_dense(), be.h.to_dense()
    offs = np.concatenate([[0], np.cumsum(sizes)])
    for s_ in range(len(sizes)):
        sl = slice(offs[s_], offs[s_ + 1])
        a = kernels.gaussian_points(pts[s_], sig[s_], 1e-2)
        eg = np.linalg.norm(ag[sl, sl] - a) / np.linalg.norm(a)
        ee = np.linalg.norm(ae[sl, sl] - a) / np.linalg.norm(a)
        print(f"\nsystem {s_}: N={sizes[s_]} sigma={sig[s_]:.3f}: recorded {eg:.1e}, adaptive {ee:.1e}", flush=True)
        assert eg <= 2 * ee + 1e-14


def test_short_refinement_finishes_eagerly():
    sizes, pts, n, b = _batch(8, 2000, 4000, 21, 4)
    sig = np.full(len(sizes), 0.3)
    dp = torch.tensor(_points(pts, sig), device="cuda")
    bt = Batch(dp, 1e-3, rtol=1e-6)  # kappa ~ 1e6: the apply alone leaves ~1e-10
    path, _ = bt.solve(sizes, b)
    assert path == "recorded"
    rec_iters = bt.h.last_refinement()[0]
    path, x = bt.solve(sizes, b)
    assert path == "replayed"
    first = bt.h.last_refinement()[1]
    target = first / 20.0
    bt.h.set_refinement(target, 10)
    path, x = bt.solve(sizes, b)
    print(f"\nrecorded with {rec_iters} refinement steps, residual {first:.1e}; target {target:.1e}: "
          f"{path}, {bt.h.last_refinement()}", flush=True)
    assert path == "redone"
    assert bt.h.last_refinement()[1] <= target
    assert _residuals(bt.h, sizes, b, x) <= 10 * target


def test_new_shape_and_moved_buffers_record_again():
    sizes, pts, n, b = _batch(6, 300, 1200, 31, 4)
    sig = np.full(len(sizes), 0.1)
    dp = torch.tensor(_points(pts, sig), device="cuda")
    bt = Batch(dp, 1e-2)
    assert bt.solve(sizes, b)[0] == "recorded"
    assert bt.solve(sizes, b)[0] == "replayed"
    # more columns: another shape
    b8 = torch.cat([b, b], 0)
    assert bt.solve(sizes, b8)[0] == "recorded"
    p, x8 = bt.solve(sizes, b8)
    assert p == "replayed"
    # a wider eager solve on the handle grows buffers the graph uses
    b16 = torch.cat([b8, b8], 0)
    x16 = torch.empty_like(b16)
    bt.h.solve_device(b16.data_ptr(), n, 16, x16.data_ptr(), n)
    torch.cuda.synchronize()
    p, x8b = bt.solve(sizes, b8)
    assert p in ("recorded", "replayed")
    assert torch.equal(x8b, x8)
    # a transposed solve drops the graph
    xt = torch.empty_like(b8)
    bt.h.solve_transpose_device(b8.data_ptr(), n, 8, xt.data_ptr(), n)
    torch.cuda.synchronize()
    p, x8c = bt.solve(sizes, b8)
    assert p == "recorded"
    assert torch.equal(x8c, x8)
