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
to it.  `roofline` is for the dominant kernel of the mixed step, from HIP
events recorded around its launches on its own stream.  `hbm_probe` is what
plain device kernels (torch's sum and copy over 2 GiB) reach on the same
GPU, next to the matvec's fraction of the 8 TB/s peak.  `cpu_baseline` is
the CPU oracle (numpy/OpenBLAS restatement of the reference algorithm) on a
bounded sample, scaled to the full size by off-diagonal area.
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


def run_mode(hmat_amd, torch, mode, a):
    """Time `a.steps` build+factor+solve steps of one precision mode."""
    n = a.n
    xs, ys = (np.arange(n) / n), ((np.arange(n) + 0.5) / n)
    dx = torch.tensor(xs, dtype=torch.float64, device="cuda")
    dy = torch.tensor(ys, dtype=torch.float64, device="cuda")
    b = torch.tensor(np.random.default_rng(0).standard_normal(n), dtype=torch.float64, device="cuda")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    prec = hmat_amd.PREC_MIXED if mode == "mixed" else hmat_amd.PREC_FP64
    o = hmat_amd.make_options(0.0, rank=a.rank, fixed_rank=True, oversample=a.oversample,
                              power_iters=a.q, exact_below=0, precision=prec,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=0x5EED)
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.set_refinement(a.rtol, 30)
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_CAUCHY, dx.data_ptr(), dy.data_ptr())
    torch.cuda.synchronize()
    times, parts, iters, resid = [], [], [], []
    for step in range(a.warmup + a.steps):
        timed = step >= a.warmup
        if timed and step == a.warmup:
            hmat_amd.profile_reset()
            hmat_amd.profile_enable(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.build_kernel_device(k, n, a.leaf)
        t1 = time.perf_counter()
        h.factor()
        t2 = time.perf_counter()
        h.solve_device(b.data_ptr(), n, 1, x.data_ptr(), n)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        it, r = h.last_refinement()
        log(f"[{mode}] step {step} ({'timed' if timed else 'warmup'}): build {t1 - t0:.3f}s "
            f"factor {t2 - t1:.3f}s solve {t3 - t2:.3f}s total {t3 - t0:.3f}s "
            f"refinement iters {it} resid {r:.2e}")
        if timed:
            times.append(t3 - t0)
            parts.append((t1 - t0, t2 - t1, t3 - t2))
            iters.append(it)
            resid.append(r)
    hmat_amd.profile_enable(False)
    prof = {}
    for name in ("sketch_gen", "sketch_lp", "sketch_lp1", "jacobi_sweeps", "jacobi_mats"):
        q = hmat_amd.profile_query(name)
        if q:
            prof[name] = {"launches": q[0], "ms": q[1], "work": q[2]}
    sweeps = prof.pop("jacobi_sweeps", None)
    mats = prof.pop("jacobi_mats", None)
    jac = round(sweeps["work"] / mats["work"], 2) if sweeps and mats and mats["work"] else None
    ranks = [max(h.level_ranks(d)) for d in range(1, h.levels() + 1)]
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
    q = hmat_amd.profile_query("matvec")
    mv = {"launches": q[0], "ms": q[1], "work": q[2]} if q else None
    return {"times": times, "parts": parts, "iters": iters, "resid": resid, "prof": prof,
            "x": x.clone(), "b": b, "dx": dx, "dy": dy, "ranks": ranks, "h": h, "matvec": mv,
            "jacobi_avg_sweeps": jac}


def true_residual(hmat_amd, torch, res, n):
    """||b - A x|| / ||b|| with the exact generated A (FP64, never stored)."""
    k = hmat_amd.make_kernel_device(hmat_amd.KERNEL_CAUCHY, res["dx"].data_ptr(), res["dy"].data_ptr())
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    hmat_amd.kernel_matvec_device(k, n, res["x"].data_ptr(), n, 1, y.data_ptr(), n)
    torch.cuda.synchronize()
    return float(torch.linalg.norm(res["b"] - y) / torch.linalg.norm(res["b"]))


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


def cpu_baseline(a):
    """Oracle build+solve of the same workload at a bounded size."""
    from oracle import hodlr, kernels, lowrank, rng
    try:
        from threadpoolctl import threadpool_info
        cores = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except Exception:
        cores = os.cpu_count() or 1
    ns = a.cpu_n
    xc, yc = kernels.cauchy_grid_points(ns)
    A = kernels.cauchy_points(xc, yc)
    o = lowrank.Options(tol=0.0, rank=a.rank, fixed_rank=True, oversample=a.oversample,
                        power_iters=a.q, exact_below=0)
    h = hodlr.HODLRMatrix(opts=o)
    b = np.random.default_rng(0).standard_normal((ns, 1))
    t0 = time.perf_counter()
    h.build_from_dense(A, a.leaf, rng.RandomStream(1))
    x = h.solve(b)
    t = time.perf_counter() - t0
    r = float(np.linalg.norm(b - h.multiply(x)) / np.linalg.norm(b))
    lev = lambda n: max(0, int(np.ceil(np.log2(n / a.leaf))))
    area = lambda n: n * n * (1.0 - 2.0 ** (-lev(n)))
    scale = area(a.n) / area(ns)
    log(f"[cpu] oracle N={ns}: {t:.2f}s resid {r:.1e}; x{scale:.0f} by off-diagonal area")
    return {"value": t * scale, "unit": "s", "cores": int(cores), "kind": "port",
            "sample": f"oracle/hodlr.py (numpy + OpenBLAS, reference algorithm incl. its "
                      f"recursive Woodbury solve) build+solve of the same Cauchy kernel at "
                      f"N={ns}, leaf {a.leaf}, rank {a.rank}, p {a.oversample}, q {a.q}: "
                      f"{t:.2f} s measured, scaled x{scale:.0f} by off-diagonal area to N={a.n}"}


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
    ap.add_argument("--rtol", type=float, default=1e-10)
    ap.add_argument("--cpu-n", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed")
    ap.add_argument("--matvec-reps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default tilings (the autotuner runs in the warm-up step otherwise)")
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
    modes = a.modes.split(",")
    res = {m: run_mode(hmat_amd, torch, m, a) for m in modes}
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
        out[f"{m}_refine_iters"] = res[m]["iters"][-1]
        out[f"{m}_resid_AH"] = res[m]["resid"][-1]
        out[f"{m}_resid_trueA"] = true_residual(hmat_amd, torch, res[m], a.n)
        out[f"{m}_max_rank_per_level"] = res[m]["ranks"]
        out[f"{m}_jacobi_avg_sweeps"] = res[m]["jacobi_avg_sweeps"]
    if "fp64" in res and "mixed" in res:
        out["speedup_mixed_vs_fp64"] = round(med["fp64"] / med["mixed"], 3)
        xf, xm = res["fp64"]["x"], res["mixed"]["x"]
        out["err_mixed_vs_fp64"] = float(torch.linalg.norm(xm - xf) / torch.linalg.norm(xf))
    # roofline of the dominant kernel of the headline step
    prof = r["prof"]
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
                           "step_share": round(p["ms"] / 1000 / a.steps / med[head], 3)}
        for k2, p2 in prof.items():
            out[f"kernel_{k2}"] = {"launches": p2["launches"], "ms_total": round(p2["ms"], 2),
                                   "tflops": round(p2["work"] / (p2["ms"] * 1e-3) / 1e12, 2)}
    mv = res[modes[0]]["matvec"] if "fp64" not in res else res["fp64"]["matvec"]
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
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
