"""C ABI: the library loads and exports every declared symbol (CPU only).

No compute call is made here; the host-side logic that needs no device
(options defaults, Omega schedule lengths, error reporting) is checked.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import hmat_amd
from oracle import hodlr as ohodlr, hss as ohss, lowrank

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hmat_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hmat_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = hmat_amd.load()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert missing == []


def test_python_binding_covers_header():
    # every declared entry point has a ctypes prototype
    lib = hmat_amd.load()
    for s in declared_symbols():
        assert getattr(lib, s).argtypes is not None, s


def test_reference_options_defaults():
    o = hmat_amd.reference_options(1e-8)
    # compression.h:72-74, hodlr_routines_omp.h:17
    assert (o.tol, o.rank, o.fixed_rank, o.oversample, o.power_iters, o.exact_below) == \
        (1e-8, 64, 0, 10, 0, 256)


@pytest.mark.parametrize("n,leaf,fixed,rank", [(4096, 512, False, 64), (1024, 128, False, 64),
                                                (1000, 60, True, 16), (4096, 64, True, 16)])
def test_omega_length_matches_oracle_schedule(n, leaf, fixed, rank):
    lib = hmat_amd.load()
    o = hmat_amd.make_options(1e-8, rank=rank, fixed_rank=fixed,
                              exact_below=0 if fixed else 256)
    oo = lowrank.Options(tol=1e-8, rank=rank, fixed_rank=fixed, exact_below=0 if fixed else 256)
    want = sum(a * b for a, b in ohodlr.omega_schedule(n, leaf, oo))
    assert lib.hmat_hodlr_omega_length(C.byref(o), n, leaf) == want


def test_create_without_device_fails_loudly():
    lib = hmat_amd.load()
    if lib.hmat_device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(hmat_amd.HmatError) as e:
        hmat_amd.HODLRMatrixOMP(1e-8)
    assert e.value.status == 2  # HMAT_ERR_DEVICE


def test_bad_options_rejected():
    lib = hmat_amd.load()
    o = hmat_amd.make_options(rank=0)
    h = C.c_void_p()
    assert lib.hmat_hodlr_create(C.byref(h), C.byref(o)) == 1  # HMAT_ERR_ARG
    assert b"rank" in lib.hmat_last_error()


def test_drop_in_driver_built_and_refuses_without_device():
    """The C++ drop-in example (include/hmat_amd_dropin.hpp) is built by build() and
    reports a missing device instead of falling back to anything."""
    import subprocess
    exe = os.path.join(ROOT, "hmat-omp_amd", "bin", "drop_in_driver")
    assert os.path.exists(exe)
    if hmat_amd.load().hmat_device_count() > 0:
        pytest.skip("a GPU is visible")
    r = subprocess.run([exe, "hodlr", "256", "64"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "no HIP device" in r.stderr
