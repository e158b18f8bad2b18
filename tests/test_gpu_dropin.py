"""A driver written with the reference's names and call sequence
(src/main_hodlr.cpp:37-107, src/main_hss.cpp:51-111) on the drop-in header
(include/hmat_amd_dropin.hpp): the reference's Matrix, Kernels::, the serial
and OpenMP class names, with the hierarchical solvers on the GPU.

Bars against the reference's own printouts (SURVEY measurements of the
unmodified reference, same matrices and tolerances):
  * HODLR, Cauchy 1/(x_i + x_j + 1) + I, N 4096, leaf 512, tol 1e-8: matvec
    error 3.06e-11 and solve error 5.75e-8 printed by the reference; the
    drop-in must be within 3x (a different Omega and x_true draw);
  * HSS, Gaussian h 0.5 + 1e-6 I, N 2048, leaf 256, tol 1e-9: matvec error
    1.91e-12 printed by the reference (within 5x); its solve error (3e2) is
    conditioning, so the solve is held to residual bars instead.  The
    reference algorithm itself (the oracle's restatement, three Omega seeds)
    leaves ||b - A x|| / ||b|| at 7.3e-4 .. 9.7e-4 on this matrix; the
    engine's factor-once push-through solve on nested bases leaves 2.4e-2
    (measured); FP64 refinement through the HSS matvec
    (hmat_hss_set_refinement) takes it to 1.5e-3 (measured), within twice
    the reference's, where the next correction no longer helps.
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "hmat-omp_amd", "bin", "drop_in_driver")


def _run(kind):
    print(f"drop_in_driver {kind}", flush=True)
    r = subprocess.run([EXE, kind], capture_output=True, text=True, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    errs = [float(e) for e in re.findall(r"Err: ([0-9.eE+-]+)", r.stdout)]
    res = [float(e) for e in re.findall(r"Resid: ([0-9.eE+-]+)", r.stdout)]
    # dense baseline, then (GEMV, Solve) for the serial and the OpenMP class
    assert len(errs) == 5 and len(res) == 2
    return errs, res


def test_hodlr_driver():
    errs, res = _run("hodlr")
    for mv, sv in ((errs[1], errs[2]), (errs[3], errs[4])):
        assert mv < 3 * 3.06e-11
        assert sv < 3 * 5.75e-8
    assert max(res) < 1e-10


def test_hss_driver():
    errs, res = _run("hss")
    for mv in (errs[1], errs[3]):
        assert mv < 5 * 1.91e-12
    assert max(res) < 5e-2


def test_hss_driver_matrix_refined():
    import numpy as np
    import hmat_amd
    from oracle import kernels
    n = 2048
    a = kernels.generate_gaussian(n, 0.5, 1e-6)
    h = hmat_amd.HSSMatrixOMP(1e-9)
    h.build_from_dense(a, 256)
    b = a @ np.random.default_rng(1).random((n, 1))
    r0 = float(np.linalg.norm(b - a @ h.solve(b)) / np.linalg.norm(b))
    h.set_refinement(2e-3, 20)
    x = h.solve(b)
    it, res = h.last_refinement()
    r1 = float(np.linalg.norm(b - a @ x) / np.linalg.norm(b))
    print(f"\nHSS on the reference driver's matrix: residual {r0:.1e} unrefined, {r1:.1e} after {it} "
          f"refinement steps (against A_H: {res:.1e})", flush=True)
    assert it >= 1 and r1 < 2e-3
