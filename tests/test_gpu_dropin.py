"""The C++ drop-in classes (include/hmat_amd.hpp) driven the way the
reference drivers drive HODLRMatrixOMP / HSSMatrixOMP (src/main_hodlr.cpp,
src/main_hss.cpp): same test matrix, b = A x_true, relative errors of the
matvec against b and of the solve against x_true.

Bars: the reference's printed errors at tol 1e-8 are about 3e-11 (matvec)
and 6e-8 (solve) for HODLR; the drop-in must be at least as accurate up to
a factor 10 margin for the different Omega draw."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_drop_in_driver_errors():
    exe = os.path.join(ROOT, "hmat-omp_amd", "bin", "drop_in_driver")
    print("drop_in_driver 4096 512 1e-8", flush=True)
    r = subprocess.run([exe, "4096", "512", "1e-8"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    errs = [float(e) for e in re.findall(r"Err: ([0-9.eE+-]+)", r.stdout)]
    assert len(errs) == 4
    hodlr_mv, hodlr_solve, hss_mv, hss_solve = errs
    assert hodlr_mv < 3e-10 and hodlr_solve < 6e-7
    assert hss_mv < 1e-8 and hss_solve < 1e-6
