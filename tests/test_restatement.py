"""The C++/OpenMP restatement (oracle/hodlr_omp.cpp, the CPU baseline) pinned
against the numpy oracle on the same Omega (the serial draw order) and the
same right-hand side: identical ranks per depth, matvec equal to rounding,
solution equal within the solve's conditioning.  CPU only; the restatement
resolves BLAS/LAPACK at run time and the test is skipped where none is
installed."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import build_restatement, hodlr as oh, kernels, lowrank, rng


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def exe():
    return build_restatement.ensure()


def _run(exe, tmp_path, args, omega, b):
    om, bf = tmp_path / "omega.bin", tmp_path / "b.bin"
    xo, yo = tmp_path / "x.bin", tmp_path / "y.bin"
    np.ascontiguousarray(omega, dtype=np.float64).tofile(om)
    np.asfortranarray(b, dtype=np.float64).T.copy().tofile(bf)  # column-major
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([exe, *args, "--omega", str(om), "--b", str(bf), "--nrhs", str(b.shape[1]),
                        "--out-x", str(xo), "--out-y", str(yo)], capture_output=True, text=True, env=env,
                       timeout=600)
    if r.returncode != 0 and "no LAPACK library found" in r.stderr:
        pytest.skip("no BLAS/LAPACK library on this machine")
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    n = b.shape[0]
    x = np.fromfile(xo).reshape(b.shape[1], n).T
    y = np.fromfile(yo).reshape(b.shape[1], n).T
    return out, x, y


@pytest.mark.parametrize("case", ["reference_options_ref_cauchy", "fixed48_q2_cauchy_points"])
def test_restatement_matches_oracle(exe, tmp_path, case):
    if case.startswith("reference"):
        n, leaf = 1024, 128
        a = kernels.generate_matrix(kernels.CAUCHY, n)
        opts = lowrank.Options.reference(1e-8)
        args = ["--kernel", "refcauchy", "--n", str(n), "--leaf", str(leaf), "--rank", "64", "--tol", "1e-8",
                "--exact-below", "256"]
    else:
        n, leaf = 2048, 256
        xc, yc = kernels.cauchy_grid_points(n)
        a = kernels.cauchy_points(xc, yc)
        opts = lowrank.Options(tol=0.0, rank=48, fixed_rank=True, oversample=10, power_iters=2, exact_below=0)
        args = ["--kernel", "cauchy", "--n", str(n), "--leaf", str(leaf), "--rank", "48", "--oversample", "10",
                "--q", "2", "--fixed", "1", "--tol", "0", "--exact-below", "0"]
    s = rng.RandomStream(17, record=True)
    ho = oh.HODLRMatrix(opts=opts)
    ho.build_from_dense(a, leaf, s)
    b = np.random.default_rng(2).standard_normal((n, 2))
    out, x, y = _run(exe, tmp_path, args, s.flat_log(), b)
    lv = ho.ranks_by_level()
    assert out["max_rank_per_depth"] == [max(max(p) for p in lv[d]) for d in range(len(lv))]
    assert out["omega_length"] == s.flat_log().size
    e_mv, e_sv = rel(y, ho.multiply_multi(b)), rel(x, ho.solve(b))
    print(f"\n{case}: matvec vs oracle {e_mv:.1e}, solve vs oracle {e_sv:.1e}, resid {out['resid']:.1e}")
    assert e_mv < 1e-13
    assert e_sv < 1e-11
    assert out["resid"] < 1e-12


def test_restatement_rejects_wrong_omega_length(exe, tmp_path):
    b = np.ones((512, 1))
    om, bf = tmp_path / "omega.bin", tmp_path / "b.bin"
    np.zeros(10).tofile(om)
    b.tofile(bf)
    r = subprocess.run([exe, "--kernel", "cauchy", "--n", "512", "--leaf", "64", "--rank", "16", "--exact-below",
                        "0", "--omega", str(om), "--b", str(bf)], capture_output=True, text=True, timeout=120)
    if "no LAPACK library found" in r.stderr:
        pytest.skip("no BLAS/LAPACK library on this machine")
    assert r.returncode == 1 and "omega holds 10 doubles" in r.stderr
