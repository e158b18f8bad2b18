#!/usr/bin/env python3
"""Benchmark: build + solve of the N = 2^18 Cauchy HODLR system to relative
residual <= 1e-10, mixed precision + FP64 refinement vs plain FP64, one GPU.

Workload (one step): A_ij = 1/(x_i - y_j), x_i = i/N, y_j = (j+1/2)/N,
N = 262144, leaf 256, fixed rank 48, oversample 10, q power iterations
(default 2, both paths), one right-hand side b ~ N(0,1) (seed 0).  A step is
build (matrix-free, entries generated on the device) + factor + solve until
||b - A_H x|| / ||b|| <= 1e-10, with A_H the FP64 compressed operator.  The
points and b are resident in HBM before the timed region.

Prints one JSON line.  `value` is the mixed-precision step time (seconds,
lower is better); the plain FP64 step time and the ratio are reported next
to it, with the true-matrix residual and the error against the FP64
solution, and the same figures at the other power-iteration counts
(`q_sweep`, q = 0 being the reference's RSVD).  Timed steps run without
kernel timing; one further step with HIP-event timing on the engine's
stream gives `roofline` (the dominant kernel of the mixed step),
`solve_roofline` (the apply-inverse) and `matvec_roofline`.  `hbm_probe` is
what plain device kernels (torch's sum and copy over 2 GiB) reach on the
same GPU.  `cpu_baseline` is the oracle's C++/OpenMP restatement of the
reference algorithm (oracle/hodlr_omp.cpp, one BLAS thread, OpenMP tasks)
measured at bounded sizes on the host cores and extrapolated to N = 2^18 by
the measured growth.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "hmat-omp_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np

METRIC = "build+solve s to 1e-10 resid, N=2^18: mixed-prec+IR vs plain FP64 GPU"
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
COPY_PROBE_BYTES = 256 << 20
# HBM bytes per dispatch from the two calibrated rocprofv3 --pmc passes
# (FETCH_SIZE x 2.000, WRITE_SIZE x 1.000, from a 256 MiB copy) recorded in
# profiles/pmc_traffic_hodlr_cauchy_n262144_mixed.txt
PMC_TRAFFIC = {"sketch_gen": 9.980e9, "sketch_lp": 10.970e9, "sketch_lp1": 4.364e9,
               "matvec": 1.0141e9 + 0.0050e9 + 1.5501e9}
PMC_SOURCE = "profiles/pmc_traffic_hodlr_cauchy_n262144_mixed.txt"
KERNEL_LABEL = {"sketch_gen": "k_sketch (FP64 MFMA, generated entries)",
                "sketch_lp": "k_sketch_lp<3> (BF16x3 MFMA, last range pass)",
                "sketch_lp1": "k_sketch_lp<1> (BF16 MFMA, power-iteration passes)"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_mode(hmat_amd, torch, mode, a, q, steps, warmup, dev):
    """Time `steps` build+factor+solve steps of one precision mode after
    `warmup` untimed ones (the first of which is the cold first call on a
    fresh handle), then one step with kernel timing."""
    n = a.n
    prec = hmat_amd.PREC_MIXED if mode == "mixed" else hmat_amd.PREC_FP64
    o = hmat_amd.make_options(0.0, rank=a.rank, fixed_rank=True, oversample=a.oversample,
                              power_iters=q, exact_below=0, precision=prec,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=0x5EED)
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.set_refinement(a.rtol, 30)
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_CAUCHY, dev["x"].data_ptr(), dev["y"].data_ptr())
    b = dev["b"]
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()

    def step():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.build_kernel_device(k, n, a.leaf)
        t1 = time.perf_counter()
        h.factor()
        t2 = time.perf_counter()
        h.solve_device(b.data_ptr(), n, 1, x.data_ptr(), n)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        return t0, t1, t2, t3

    times, parts, first = [], [], None
    for i in range(warmup + steps):
        t0, t1, t2, t3 = step()
        it, r = h.last_refinement()
        timed = i >= warmup
        log(f"[{mode} q={q}] step {i} ({'timed' if timed else 'warmup'}): build {t1 - t0:.3f}s "
            f"factor {t2 - t1:.3f}s solve {t3 - t2:.3f}s total {t3 - t0:.3f}s "
            f"refinement iters {it} resid {r:.2e}")
        if i == 0:
            first = t3 - t0
        if timed:
            times.append(t3 - t0)
            parts.append((t1 - t0, t2 - t1, t3 - t2))
    iters, resid = h.last_refinement()
    out = {"times": times, "parts": parts, "first_call_s": first, "iters": iters, "resid": resid,
           "x": x.clone(), "ranks": [max(h.level_ranks(d)) for d in range(1, h.levels() + 1)]}
    if not a.profile:
        return out, h
    # one more step with HIP-event timing of the named launches
    hmat_amd.profile_reset()
    hmat_amd.profile_enable(True)
    step()
    # further applies of the factored inverse for its own average
    for _ in range(a.solve_reps):
        h.solve_device(b.data_ptr(), n, 1, x.data_ptr(), n)
    torch.cuda.synchronize()
    hmat_amd.profile_enable(False)
    prof = {}
    for name in ("sketch_gen", "sketch_lp", "sketch_lp1", "jacobi_sweeps", "jacobi_mats", "solve_apply"):
        pq = hmat_amd.profile_query(name)
        if pq:
            prof[name] = {"launches": pq[0], "ms": pq[1], "work": pq[2]}
    sweeps, mats = prof.pop("jacobi_sweeps", None), prof.pop("jacobi_mats", None)
    out["jacobi_avg_sweeps"] = round(sweeps["work"] / mats["work"], 2) if sweeps and mats and mats["work"] else None
    out["solve"] = prof.pop("solve_apply", None)
    out["prof"] = prof
    # matvec y = A_H x of the built operator (FP64 generators), 1 RHS
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    for _ in range(3):
        h.multiply_device(b.data_ptr(), n, 1, y.data_ptr(), n)
    torch.cuda.synchronize()
    hmat_amd.profile_reset()
    hmat_amd.profile_enable(True)
    for _ in range(a.matvec_reps):
        h.multiply_device(b.data_ptr(), n, 1, y.data_ptr(), n)
    hmat_amd.profile_enable(False)
    mq = hmat_amd.profile_query("matvec")
    out["matvec"] = {"launches": mq[0], "ms": mq[1], "work": mq[2]} if mq else None
    return out, h


def batch_leg(hmat_amd, torch, steps, warmup, profile=True):
    """The batched workload: 256 independent Gaussian-kernel systems,
    A_s = exp(-(p_i - p_j)^2 / (2 sigma_s^2)) + 1e-2 delta_ij, N_s
    log-uniform in [1024, 16384], sigma_s ~ U[0.05, 0.3], sorted U[0,1]
    points (seeded), leaf 64, fixed rank 24, p 10, q 1, RHS counts uniform in
    [1, 64] (zero-padded to 64 columns).  One step = build of the forest +
    factor + FP64 solve with refinement to 1e-11 per column; the worst
    per-system, per-column residual ||b - A_H x|| / ||b|| is reported."""
    rng = np.random.default_rng(2024)
    nsys, lam = 256, 1e-2
    sizes = np.exp(rng.uniform(np.log(1024), np.log(16384), nsys)).astype(np.int64)
    sigmas = rng.uniform(0.05, 0.3, nsys)
    nrhs = rng.integers(1, 65, nsys)
    pts = [np.sort(rng.random(int(s))) for s in sizes]
    n = int(sizes.sum())
    cols = int(nrhs.max())
    offs = np.concatenate([[0], np.cumsum(sizes)])
    bh = np.zeros((cols, n))  # column-major n x cols
    for s in range(nsys):
        bh[:nrhs[s], offs[s]:offs[s + 1]] = rng.standard_normal((nrhs[s], sizes[s]))
    dp = torch.tensor(np.concatenate([p / sg for p, sg in zip(pts, sigmas)]), device="cuda")
    db = torch.tensor(bh, device="cuda")
    dx = torch.zeros_like(db)
    o = hmat_amd.make_options(0.0, rank=24, fixed_rank=True, oversample=10, power_iters=1, exact_below=0,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=99)
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.set_refinement(1e-11, 10)
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_GAUSSIAN, dp.data_ptr(), p0=1.0, p1=lam)
    szl = [int(s) for s in sizes]

    def step():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.build_kernel_batch_device(k, szl, 64)
        t1 = time.perf_counter()
        h.factor()
        t2 = time.perf_counter()
        h.solve_device(db.data_ptr(), n, cols, dx.data_ptr(), n)
        torch.cuda.synchronize()
        return t0, t1, t2, time.perf_counter()

    times, parts = [], []
    for i in range(warmup + steps):
        t0, t1, t2, t3 = step()
        log(f"[batch] step {i} ({'timed' if i >= warmup else 'warmup'}): build {t1 - t0:.3f}s "
            f"factor {t2 - t1:.3f}s solve {t3 - t2:.3f}s total {t3 - t0:.3f}s "
            f"refinement {h.last_refinement()}")
        if i >= warmup:
            times.append(t3 - t0)
            parts.append((t1 - t0, t2 - t1, t3 - t2))

    # the same step as one recorded graph (hmat_hodlr_batch_solve_device):
    # the first call records, later calls replay with new widths written into
    # the points in place (sigma_s rescaled by 1 + 1e-3 i)
    hg = hmat_amd.HODLRMatrixOMP(opts=o)
    hg.set_refinement(1e-11, 10)
    dxg = torch.zeros_like(db)
    gtimes, gpaths = [], []
    for i in range(warmup + 1 + steps):
        dp.copy_(torch.tensor(np.concatenate([p / (sg * (1 + 1e-3 * i)) for p, sg in zip(pts, sigmas)]),
                              device="cuda"))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        path = hg.batch_solve_device(k, szl, 64, db.data_ptr(), n, cols, dxg.data_ptr(), n, graph=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        log(f"[batch graph] call {i}: {path} {t1 - t0:.3f}s refinement {hg.last_refinement()}")
        gpaths.append(path)
        if i >= warmup + 1:
            gtimes.append(t1 - t0)
    dyg = torch.empty_like(db)
    hg.multiply_device(dxg.data_ptr(), n, cols, dyg.data_ptr(), n)
    torch.cuda.synchronize()
    rg = (db - dyg).cpu().numpy()
    worst_g = 0.0
    for s in range(nsys):
        sl = slice(offs[s], offs[s + 1])
        for j in range(int(nrhs[s])):
            worst_g = max(worst_g, float(np.linalg.norm(rg[j, sl]) / np.linalg.norm(bh[j, sl])))
    del hg
    dy = torch.empty_like(db)
    h.multiply_device(dx.data_ptr(), n, cols, dy.data_ptr(), n)
    torch.cuda.synchronize()
    r = (db - dy).cpu().numpy()
    worst = 0.0
    for s in range(nsys):
        sl = slice(offs[s], offs[s + 1])
        for j in range(int(nrhs[s])):
            worst = max(worst, float(np.linalg.norm(r[j, sl]) / np.linalg.norm(bh[j, sl])))
    out = {"workload": "256 Gaussian HODLR systems, N log-uniform [1024,16384], sigma U[0.05,0.3], "
                       "lambda 1e-2, leaf 64, rank 24, p 10, q 1, RHS U[1,64] (64 columns), FP64 + IR 1e-11",
           "sum_N": n, "rhs_total": int(nrhs.sum()), "columns": cols,
           "s_per_step": round(statistics.median(times), 5),
           "build_factor_solve_s": [round(statistics.median(p[i] for p in parts), 5) for i in range(3)],
           "refine_iters": h.last_refinement()[0], "worst_system_resid": worst,
           "graph": {"s_per_step": round(statistics.median(gtimes), 5), "paths": gpaths,
                     "worst_system_resid": worst_g,
                     "note": "build + factor + solve replayed from one HIP graph; widths change per call"}}
    if profile:
        hmat_amd.profile_reset()
        hmat_amd.profile_enable(True)
        step()
        hmat_amd.profile_enable(False)
        prof = {}
        for name in ("sketch_gen", "orth_gram", "orth_trsm", "solve_apply", "factor"):
            pq = hmat_amd.profile_query(name)
            if pq:
                prof[name] = {"launches": pq[0], "ms": round(pq[1], 3), "work": pq[2]}
        out["kernels"] = prof
        sk = prof.get("sketch_gen")
        if sk and sk["ms"] > 0:
            # the batch's dominant kernel: the FP64 generated-entry sketches
            # (Gaussian entries), 2 k (2 + 2 q) sum(m n) flops per step
            achieved = sk["work"] / (sk["ms"] * 1e-3) / 1e12
            out["roofline"] = {"kernel": "k_sketch (FP64 MFMA, generated Gaussian entries)", "bound": "mfma",
                               "achieved": round(achieved, 2), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                               "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "launches": sk["launches"],
                               "step_share": round(sk["ms"] / 1000 / statistics.median(times), 3)}
    del h
    return out


def hss_leg(hmat_amd, torch, reps=10):
    """The HSS + HODLR workload: Gaussian kernel exp(-(p_i - p_j)^2 / (2 sigma^2))
    + lambda delta_ij, sigma 0.05, lambda 1e-2, N = 65536 sorted U[0,1] points
    (seeded), leaf 128, fixed rank 32, p 10, q 2, 16 right-hand sides ~ N(0,1),
    FP64.  Nested HSS and HODLR on the same matrix: build, factor, solve and
    matvec times, the apply-inverse and matvec HBM rates (algorithmic bytes
    over HIP-event times of `reps` calls each), residuals against A_H."""
    n, sigma, lam, leaf, rank, nrhs = 65536, 0.05, 1e-2, 128, 32, 16
    rng = np.random.default_rng(5)
    pts = np.sort(rng.random(n))
    db = torch.tensor(rng.standard_normal((nrhs, n)), device="cuda")  # column-major n x nrhs
    dx = torch.empty_like(db)
    dy = torch.empty_like(db)
    dp = torch.tensor(pts, device="cuda")
    o = hmat_amd.make_options(0.0, rank=rank, fixed_rank=True, oversample=10, power_iters=2, exact_below=0,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=1)
    out = {"workload": f"Gaussian sigma {sigma} lambda {lam}, N={n} sorted U[0,1], leaf {leaf}, rank {rank}, "
                       f"p 10, q 2, {nrhs} RHS, FP64 (nested HSS from N-wide leaf blocks, nested HSS from "
                       f"one sample pass per side (q does not apply), and HODLR)"}
    for name in ("hss", "hss_sampled", "hodlr"):
        if name == "hss":
            h = hmat_amd.HSSMatrixOMP(opts=o, nested=True)
            k = hmat_amd.make_kernel(hmat_amd.KERNEL_GAUSSIAN, pts, p0=sigma, p1=lam)
        elif name == "hss_sampled":
            h = hmat_amd.HSSMatrixOMP(opts=o, nested=hmat_amd.HSS_SAMPLED)
            k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_GAUSSIAN, dp.data_ptr(), p0=sigma, p1=lam)
        else:
            h = hmat_amd.HODLRMatrixOMP(opts=o)
            k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_GAUSSIAN, dp.data_ptr(), p0=sigma, p1=lam)
        r = {}
        for it in range(2):  # the first build on a fresh handle allocates
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if name == "hss":
                h.build_kernel(k, n, leaf)
            else:
                h.build_kernel_device(k, n, leaf)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            h.factor()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            h.solve_device(db.data_ptr(), n, nrhs, dx.data_ptr(), n)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
        r.update(build_s=round(t1 - t0, 4), factor_s=round(t2 - t1, 4), solve_s=round(t3 - t2, 5))
        h.multiply_device(dx.data_ptr(), n, nrhs, dy.data_ptr(), n)
        torch.cuda.synchronize()
        r["resid_AH"] = float(torch.linalg.norm(db - dy) / torch.linalg.norm(db))
        if name != "hodlr":
            r["root_rank"] = h.root_rank()
        hmat_amd.profile_reset()
        hmat_amd.profile_enable(True)
        for _ in range(reps):
            h.multiply_device(db.data_ptr(), n, nrhs, dy.data_ptr(), n)
        for _ in range(reps):
            h.solve_device(db.data_ptr(), n, nrhs, dx.data_ptr(), n)
        torch.cuda.synchronize()
        hmat_amd.profile_enable(False)
        for key, pname in (("matvec", "matvec" if name == "hodlr" else "hss_matvec"),
                           ("solve", "solve_apply" if name == "hodlr" else "hss_solve_apply")):
            pq = hmat_amd.profile_query(pname)
            if pq and pq[0]:
                ms = pq[1] / pq[0]
                gbs = pq[2] / pq[0] / (ms * 1e-3) / 1e9
                r[key] = {"avg_ms": round(ms, 4), "bytes": pq[2] / pq[0], "GBs": round(gbs, 1),
                          "frac_hbm": round(gbs / PEAK_HBM_GBS, 4), "calls": pq[0]}
        log(f"[{name}] {r}")
        out[name] = r
        del h
    return out


def hss_cauchy_leg(hmat_amd, torch, dev, n, leaf, rank, p, rtol, reps=10):
    """Nested HSS (one sample pass per side) on the precision benchmark's
    system: 1/(x_i - y_j) on the grid, N = 2^18, leaf 256, fixed rank 48,
    p 10, FP64 with refinement through the HSS matvec to rtol.  Build,
    factor and solve times (second call on the handle), the residual
    against A_H and the true A, and the matvec / apply-inverse HBM rates
    (algorithmic bytes over HIP-event times of `reps` calls each)."""
    o = hmat_amd.make_options(0.0, rank=rank, fixed_rank=True, oversample=p, power_iters=0, exact_below=0,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=0x5EED)
    h = hmat_amd.HSSMatrixOMP(opts=o, nested=hmat_amd.HSS_SAMPLED)
    h.set_refinement(rtol, 10)
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_CAUCHY, dev["x"].data_ptr(), dev["y"].data_ptr())
    b = dev["b"]
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    r = {"workload": f"HSS (sampled nested bases) Cauchy 1/(x_i-y_j), N={n}, leaf {leaf}, rank {rank}, p {p}, "
                     f"1 RHS, FP64 + refinement to {rtol}"}
    for it in range(2):  # the first build on a fresh handle allocates
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.build_kernel_device(k, n, leaf)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        h.factor()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        h.solve_device(b.data_ptr(), n, 1, x.data_ptr(), n)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        if it == 0:
            r["first_call_s"] = round(t3 - t0, 4)
    r.update(build_s=round(t1 - t0, 4), factor_s=round(t2 - t1, 4), solve_s=round(t3 - t2, 5),
             total_s=round(t3 - t0, 4), root_rank=h.root_rank())
    its, res = h.last_refinement()
    r.update(refine_iters=its, resid_AH=res, resid_trueA=true_residual(hmat_amd, torch, x, dev, n))
    hmat_amd.profile_reset()
    hmat_amd.profile_enable(True)
    for _ in range(reps):
        h.multiply_device(b.data_ptr(), n, 1, y.data_ptr(), n)
    h.set_refinement(0.0)
    for _ in range(reps):
        h.solve_device(b.data_ptr(), n, 1, x.data_ptr(), n)
    torch.cuda.synchronize()
    hmat_amd.profile_enable(False)
    for key, pname in (("matvec", "hss_matvec"), ("solve", "hss_solve_apply")):
        pq = hmat_amd.profile_query(pname)
        if pq and pq[0]:
            ms = pq[1] / pq[0]
            gbs = pq[2] / pq[0] / (ms * 1e-3) / 1e9
            r[key] = {"avg_ms": round(ms, 4), "bytes": pq[2] / pq[0], "GBs": round(gbs, 1),
                      "frac_hbm": round(gbs / PEAK_HBM_GBS, 4), "calls": pq[0]}
    # ULV factorisation of the same operator (orthogonal eliminations,
    # backward stable): factor and solve times (second call), residuals of
    # the unrefined solve, and the apply-inverse rate
    h.set_solver(hmat_amd.HSS_ULV)
    u = {}
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.factor()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        h.solve_device(b.data_ptr(), n, 1, x.data_ptr(), n)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
    u.update(factor_s=round(t1 - t0, 4), solve_s=round(t2 - t1, 5))
    h.multiply_device(x.data_ptr(), n, 1, y.data_ptr(), n)
    torch.cuda.synchronize()
    u.update(resid_AH=float(torch.linalg.norm(b - y) / torch.linalg.norm(b)),
             resid_trueA=true_residual(hmat_amd, torch, x, dev, n))
    hmat_amd.profile_reset()
    hmat_amd.profile_enable(True)
    for _ in range(reps):
        h.solve_device(b.data_ptr(), n, 1, x.data_ptr(), n)
    torch.cuda.synchronize()
    hmat_amd.profile_enable(False)
    pq = hmat_amd.profile_query("hss_ulv_apply")
    if pq and pq[0]:
        ms = pq[1] / pq[0]
        gbs = pq[2] / pq[0] / (ms * 1e-3) / 1e9
        u["solve"] = {"avg_ms": round(ms, 4), "bytes": pq[2] / pq[0], "GBs": round(gbs, 1),
                      "frac_hbm": round(gbs / PEAK_HBM_GBS, 4), "calls": pq[0]}
    r["ulv"] = u
    log(f"[hss_cauchy] {r}")
    del h
    return r


def true_residual(hmat_amd, torch, x, dev, n):
    """||b - A x|| / ||b|| with the exact generated A (FP64, never stored)."""
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_CAUCHY, dev["x"].data_ptr(), dev["y"].data_ptr())
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    hmat_amd.kernel_matvec_device(k, n, x.data_ptr(), n, 1, y.data_ptr(), n)
    torch.cuda.synchronize()
    return float(torch.linalg.norm(dev["b"] - y) / torch.linalg.norm(dev["b"]))


def hbm_probe(torch, nbytes=1 << 31, reps=5):
    """What plain device kernels reach on this GPU's HBM: torch's sum (a pure
    read of `nbytes`) and copy (`nbytes` read + `nbytes` written), timed with
    events on torch's stream after one untimed call each."""
    src = torch.rand(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)

    def ms(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    t_read, t_copy = ms(lambda: src.sum()), ms(lambda: dst.copy_(src))
    del src, dst
    torch.cuda.empty_cache()
    return {"bytes": nbytes, "read_GBs": round(nbytes / (t_read * 1e-3) / 1e9, 1),
            "copy_GBs": round(2 * nbytes / (t_copy * 1e-3) / 1e9, 1)}


def _lapack_env():
    """Environment for the restatement: MKL's single dynamic library when the
    image has it, else scipy's bundled OpenBLAS (symbols with a scipy_
    prefix, which the restatement also accepts)."""
    import glob
    env = dict(os.environ)
    if not os.path.exists("/opt/conda/lib/libmkl_rt.so.1") and "HMAT_LAPACK" not in env:
        import scipy
        libs = sorted(glob.glob(os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs",
                                             "libscipy_openblas*.so")))
        if libs:
            env["HMAT_LAPACK"] = libs[0]
    return env


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(a):
    """The oracle's C++/OpenMP restatement of HODLRMatrixOMP (oracle/hodlr_omp.cpp):
    build + recursive Woodbury solve of the same Cauchy workload (leaf, rank,
    oversample, q) at bounded sizes, OpenMP tasks over `threads` host cores,
    one BLAS thread (run_tests.sh:7-11).  The full size is extrapolated from
    the two largest measured sizes by their measured growth exponent."""
    import subprocess
    from oracle import build_restatement
    exe = build_restatement.ensure()
    threads = a.cpu_threads or min(16, os.cpu_count() or 1)
    env = _lapack_env()
    env["OMP_NUM_THREADS"] = str(threads)
    meas = []
    for ns in [int(v) for v in a.cpu_sizes.split(",")]:
        cmd = [exe, "--kernel", "cauchy", "--n", str(ns), "--leaf", str(a.leaf), "--rank", str(a.rank),
               "--oversample", str(a.oversample), "--q", str(a.q), "--fixed", "1", "--tol", "0",
               "--exact-below", "0", "--seed", "1"]
        r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=900)
        if r.returncode != 0:
            log(f"[cpu] restatement failed at N={ns}: {r.stderr.strip()}")
            return None
        d = json.loads(r.stdout)
        meas.append({"n": ns, "build_s": round(d["build_s"], 4), "solve_s": round(d["solve_s"], 4),
                     "total_s": round(d["build_s"] + d["solve_s"], 4), "resid_AH": d["resid"]})
        log(f"[cpu] restatement N={ns}: build {d['build_s']:.3f}s solve {d['solve_s']:.3f}s "
            f"resid {d['resid']:.1e} ({threads} threads, {d['lapack']})")
    lev = lambda n: max(0, int(np.ceil(np.log2(n / a.leaf))))
    area = lambda n: n * n * (1.0 - 2.0 ** (-lev(n)))
    m1, m2 = meas[-2], meas[-1]
    expo = float(np.log(m2["total_s"] / m1["total_s"]) / np.log(m2["n"] / m1["n"]))
    value = m2["total_s"] * (a.n / m2["n"]) ** expo
    area_pred = m1["total_s"] * area(m2["n"]) / area(m1["n"])
    return {"value": round(value, 1), "unit": "s", "cores": int(threads), "kind": "port",
            "sample": (f"oracle/hodlr_omp.cpp (C++/OpenMP restatement of HODLRMatrixOMP: tasks above 512 rows, "
                       f"recursive Woodbury solve, one BLAS thread) build+solve of the same Cauchy workload "
                       f"(leaf {a.leaf}, rank {a.rank}, p {a.oversample}, q {a.q}) measured at N="
                       f"{','.join(str(m['n']) for m in meas)} on {threads} threads of {cpu_model()}; "
                       f"value = the N={m2['n']} time x (N/{m2['n']})^{expo:.2f}, the growth exponent measured "
                       f"between the two largest sizes"),
            "measured": meas, "growth_exponent": round(expo, 3),
            "area_scaling_check": {"predicted_s": round(area_pred, 4), "measured_s": m2["total_s"],
                                   "measured_over_predicted": round(m2["total_s"] / area_pred, 3)},
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 18)
    ap.add_argument("--leaf", type=int, default=256)
    ap.add_argument("--rank", type=int, default=48)
    ap.add_argument("--oversample", type=int, default=10)
    ap.add_argument("--q", type=int, default=2)
    ap.add_argument("--q-sweep", default="0,1",
                    help="further power-iteration counts measured beside the headline q ('' for none)")
    ap.add_argument("--sweep-steps", type=int, default=3)
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--cpu-sizes", default="4096,8192,16384")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed",
                    help="precision legs of the headline workload ('' runs only the batch and HSS legs)")
    ap.add_argument("--matvec-reps", type=int, default=20)
    ap.add_argument("--solve-reps", type=int, default=10)
    ap.add_argument("--no-profile", dest="profile", action="store_false",
                    help="skip the kernel-timed step (rooflines)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tilings (the autotuner runs in the warm-up step otherwise)")
    ap.add_argument("--batch-steps", type=int, default=3,
                    help="timed steps of the batched leg (256 systems); 0 skips it")
    ap.add_argument("--no-hss", action="store_true",
                    help="skip the HSS legs (HSS + HODLR at N=65536 with 16 RHS; sampled HSS on the headline system)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 256 MiB device copy first (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); the engine has no "
              f"multi-GPU path.", file=sys.stderr)
        sys.exit(2)
    import torch
    import hmat_amd
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if a.copy_probe:
        # a plain device copy of known size: its FETCH_SIZE / WRITE_SIZE counters
        # calibrate the per-dispatch HBM traffic of the engine's kernels
        src = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if not a.no_autotune:
        hmat_amd.autotune_enable(True)
    n = a.n
    dev = {"x": torch.tensor(np.arange(n) / n, dtype=torch.float64, device="cuda"),
           "y": torch.tensor((np.arange(n) + 0.5) / n, dtype=torch.float64, device="cuda"),
           "b": torch.tensor(np.random.default_rng(0).standard_normal(n), dtype=torch.float64, device="cuda")}
    modes = [m for m in a.modes.split(",") if m]
    if not modes:
        # only the batch and HSS legs (kernel profiles of those workloads)
        out = {"metric": METRIC, "value": None, "unit": "s", "n_gpus": 1,
               "note": "headline legs skipped (--modes '')"}
        if a.batch_steps > 0:
            out["batch"] = batch_leg(hmat_amd, torch, a.batch_steps, 1, a.profile)
        if not a.no_hss:
            out["hss_hodlr_n65536"] = hss_leg(hmat_amd, torch)
            out["hss_cauchy_n262144"] = hss_cauchy_leg(hmat_amd, torch, dev, n, a.leaf, a.rank, a.oversample, a.rtol)
        print(json.dumps(out), flush=True)
        return
    res = {}
    for m in modes:
        r, h = run_mode(hmat_amd, torch, m, a, a.q, a.steps, a.warmup, dev)
        del h
        res[m] = r
    med = {m: statistics.median(r["times"]) for m, r in res.items()}
    head = "mixed" if "mixed" in res else modes[0]
    r = res[head]
    out = {
        "metric": METRIC,
        "value": round(med[head], 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1000 * med[head], 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": ("bf16 power passes, bf16x3 last range pass, f64 projection, f32 factor, f64 IR"
                  if head == "mixed" else "f64"),
        "data": "synthetic (deterministic Cauchy grid, b ~ N(0,1) seed 0)",
        "config": {"workload": f"HODLR Cauchy 1/(x_i-y_j), N={a.n}, leaf {a.leaf}, rank {a.rank}, "
                               f"p {a.oversample}, q {a.q}, 1 RHS, resid<={a.rtol:g}",
                   "N": a.n, "leaf": a.leaf, "rank": a.rank, "oversample": a.oversample,
                   "power_iters": a.q, "rtol": a.rtol, "parallelism": "single GPU",
                   "autotune": not a.no_autotune},
    }
    for m in modes:
        out[f"{m}_s"] = round(med[m], 5)
        bp = [statistics.median(p[i] for p in res[m]["parts"]) for i in range(3)]
        out[f"{m}_build_factor_solve_s"] = [round(v, 5) for v in bp]
        out[f"{m}_first_call_s"] = round(res[m]["first_call_s"], 4)
        out[f"{m}_refine_iters"] = res[m]["iters"]
        out[f"{m}_resid_AH"] = res[m]["resid"]
        out[f"{m}_resid_trueA"] = true_residual(hmat_amd, torch, res[m]["x"], dev, n)
        out[f"{m}_max_rank_per_level"] = res[m]["ranks"]
        out[f"{m}_jacobi_avg_sweeps"] = res[m].get("jacobi_avg_sweeps")
    if "fp64" in res and "mixed" in res:
        out["speedup_mixed_vs_fp64"] = round(med["fp64"] / med["mixed"], 3)
        xf, xm = res["fp64"]["x"], res["mixed"]["x"]
        out["err_mixed_vs_fp64"] = float(torch.linalg.norm(xm - xf) / torch.linalg.norm(xf))
    # the same workload at the other power-iteration counts (q = 0: the
    # reference's RSVD, compression.h:77-90, has no power iterations)
    sweep = []
    for qv in [int(v) for v in a.q_sweep.split(",") if v.strip() != ""]:
        if qv == a.q:
            continue
        ent = {"q": qv}
        xs = {}
        for m in modes:
            rr, h = run_mode(hmat_amd, torch, m, a, qv, a.sweep_steps, 1, dev)
            del h
            ent[f"{m}_s"] = round(statistics.median(rr["times"]), 5)
            ent[f"{m}_first_call_s"] = round(rr["first_call_s"], 4)
            ent[f"{m}_resid_AH"] = rr["resid"]
            ent[f"{m}_resid_trueA"] = true_residual(hmat_amd, torch, rr["x"], dev, n)
            if rr.get("prof", {}).get("sketch_gen"):
                p = rr["prof"]["sketch_gen"]
                ent[f"{m}_fp64_sketch_ms"] = round(p["ms"], 2)
            xs[m] = rr["x"]
        if "fp64" in xs and "mixed" in xs:
            ent["speedup_mixed_vs_fp64"] = round(ent["fp64_s"] / ent["mixed_s"], 3)
            ent["err_mixed_vs_fp64"] = float(torch.linalg.norm(xs["mixed"] - xs["fp64"]) /
                                             torch.linalg.norm(xs["fp64"]))
        sweep.append(ent)
        torch.cuda.empty_cache()
    if sweep:
        out["q_sweep"] = sweep
    # roofline of the dominant kernel of the headline step
    prof = r.get("prof", {})
    name = max(prof, key=lambda k: prof[k]["ms"]) if prof else None
    if name:
        p = prof[name]
        achieved = p["work"] / (p["ms"] * 1e-3) / 1e12
        peak = PEAK_FP64_TFLOPS if name == "sketch_gen" else PEAK_BF16_TFLOPS
        out["roofline"] = {"kernel": KERNEL_LABEL[name], "bound": "mfma",
                           "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                           "frac": round(achieved / peak, 4),
                           "traffic": PMC_TRAFFIC.get(name) if a.n == 1 << 18 else None,
                           "traffic_source": PMC_SOURCE,
                           "launches": p["launches"], "avg_launch_ms": round(p["ms"] / p["launches"], 4),
                           "step_share": round(p["ms"] / 1000 / med[head], 3)}
        for k2, p2 in prof.items():
            out[f"kernel_{k2}"] = {"launches": p2["launches"], "ms_total": round(p2["ms"], 2),
                                   "tflops": round(p2["work"] / (p2["ms"] * 1e-3) / 1e12, 2)}
    sr = {}
    for m in modes:
        sv = res[m].get("solve")
        if sv:
            avg = sv["ms"] / sv["launches"]
            gbs = sv["work"] / sv["launches"] / (avg * 1e-3) / 1e9
            sr[m] = {"kernel": ("apply-inverse of the FP64 factorisation (refined leaf and capacitance solves)"
                                if m == "fp64" else "apply-inverse of the FP32 factorisation"),
                     "bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(gbs / PEAK_HBM_GBS, 4), "bytes_per_launch": sv["work"] / sv["launches"],
                     "avg_launch_ms": round(avg, 4), "launches": sv["launches"]}
    if sr:
        out["solve_roofline"] = sr
    mv = res["fp64"].get("matvec") if "fp64" in res else res[modes[0]].get("matvec")
    if mv:
        avg = mv["ms"] / mv["launches"]
        gbs = mv["work"] / mv["launches"] / (avg * 1e-3) / 1e9
        out["matvec_roofline"] = {"kernel": "fused HODLR matvec (project + reduce + expand), FP64, 1 RHS",
                                  "bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS,
                                  "unit": "GB/s", "frac": round(gbs / PEAK_HBM_GBS, 4),
                                  "traffic": PMC_TRAFFIC["matvec"] if a.n == 1 << 18 else None,
                                  "traffic_source": PMC_SOURCE,
                                  "bytes_per_launch": mv["work"] / mv["launches"],
                                  "avg_launch_ms": round(avg, 4), "launches": mv["launches"]}
        probe = hbm_probe(torch)
        out["hbm_probe"] = probe
        out["matvec_roofline"]["frac_of_probe_read"] = round(gbs / probe["read_GBs"], 4)
    if a.batch_steps > 0:
        out["batch"] = batch_leg(hmat_amd, torch, a.batch_steps, 1, a.profile)
    if not a.no_hss:
        out["hss_hodlr_n65536"] = hss_leg(hmat_amd, torch)
        out["hss_cauchy_n262144"] = hss_cauchy_leg(hmat_amd, torch, dev, n, a.leaf, a.rank, a.oversample, a.rtol)
    if not a.no_cpu:
        cb = cpu_baseline(a)
        if cb:
            out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
