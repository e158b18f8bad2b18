"""The allocation-free matvec entry point against hmat_hodlr_matvec_device.

Two calls in a row on one GPU share the fixed scratch array: a larger
operator first (four columns), then a smaller one (one column), each
compared with the allocating entry point (the same kernels, so bit for bit
up to four columns; to rounding beyond, where the allocating entry point
takes its segmented path).  An operator whose partials exceed the array is
rejected with HMAT_ERR_ARG and the allocating entry point still serves it.
"""
import numpy as np
import pytest
import torch

import hmat_amd
from oracle import kernels

pytestmark = pytest.mark.gpu


def _handle(n, leaf, rank):
    xc, yc = kernels.cauchy_grid_points(n)
    o = hmat_amd.make_options(0.0, rank=rank, fixed_rank=True, oversample=10, power_iters=1, exact_below=0,
                              omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=3)
    h = hmat_amd.HODLRMatrixOMP(opts=o)
    h.build_kernel(hmat_amd.make_kernel(hmat_amd.KERNEL_CAUCHY, xc, yc), n, leaf)
    return h


def _both(h, n, nrhs, seed):
    x = torch.tensor(np.random.default_rng(seed).standard_normal((nrhs, n)), device="cuda")  # column-major n x nrhs
    y0 = torch.full_like(x, np.nan)
    y1 = torch.full_like(x, np.nan)
    torch.cuda.synchronize()
    h.multiply_device(x.data_ptr(), n, nrhs, y0.data_ptr(), n)
    h.multiply_device_noalloc(x.data_ptr(), n, nrhs, y1.data_ptr(), n)
    torch.cuda.synchronize()
    return y0.cpu().numpy(), y1.cpu().numpy()


def test_noalloc_matches_allocating_entry_point():
    print("\nno-alloc: N=16384 leaf 128 rank 48, 4 columns", flush=True)
    big = _handle(16384, 128, 48)
    y0, y1 = _both(big, 16384, 4, 1)
    assert np.array_equal(y0, y1)
    print("no-alloc: N=2048 leaf 64 rank 16, 1 column (after the larger call)", flush=True)
    small = _handle(2048, 64, 16)
    y0, y1 = _both(small, 2048, 1, 2)
    assert np.array_equal(y0, y1)
    # the larger operator again, three columns (passes of 2 + 1)
    y0, y1 = _both(big, 16384, 3, 3)
    assert np.array_equal(y0, y1)
    # beyond four columns the allocating entry point takes its segmented path
    y0, y1 = _both(small, 2048, 7, 4)
    assert np.linalg.norm(y1 - y0) <= 1e-14 * np.linalg.norm(y0)


def test_noalloc_rejects_operator_beyond_scratch():
    cap = hmat_amd.load().hmat_hodlr_matvec_noalloc_capacity()
    assert cap == 4 << 20
    # 4096 leaves x 12 depths x rank 24 partials: 1.18 Mi doubles per column,
    # 4.7 Mi at four columns (above the 4 Mi array), within it at one
    print("no-alloc: N=131072 leaf 32 rank 24, 4 columns (rejected) then 1", flush=True)
    h = _handle(1 << 17, 32, 24)
    with pytest.raises(hmat_amd.HmatError) as e:
        _both(h, 1 << 17, 4, 5)
    assert e.value.status == 1
    y0, y1 = _both(h, 1 << 17, 1, 6)
    assert np.array_equal(y0, y1)
