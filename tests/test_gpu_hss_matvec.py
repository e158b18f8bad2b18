"""The one-launch HSS matvec against the dense reconstruction of the same
operator (hss_routines_omp.h:154-230).

Up to 16 columns run as one persistent launch (leaf products and g = V^T x,
both sweeps as a queue of node tasks, then y += U f), wider calls as groups
of 16, 8, 4, 2 and 1 columns.  Every column of every group width equals the
dense reconstruction's product, and the one-column product of a column
equals that column of a wider product.  The profiler's one-launch counter
shows which calls took the persistent kernel; operators beyond its limits
(ranks above 64) keep the per-depth launches.
"""
import numpy as np
import pytest

import hmat_amd
from oracle import kernels

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def cauchy(n):
    x = np.arange(n) / n
    return 1.0 / (x[:, None] - (x[None, :] + 0.5 / n))


def opts(rank, p, tol=0.0, fixed=True):
    return hmat_amd.make_options(tol, rank=rank, fixed_rank=fixed, oversample=p, power_iters=0, exact_below=0,
                                 omega_dist=hmat_amd.OMEGA_GAUSSIAN, seed=21)


CASES = {
    # the precision benchmark's kernel at leaf 256, fixed rank 48 (sampled bases)
    "sampled_cauchy": (lambda n: cauchy(n), 4096, 256, hmat_amd.HSS_SAMPLED, opts(48, 10)),
    # ragged tree of 100-row leaves, Gaussian kernel, rank 32
    "sampled_gaussian_ragged": (lambda n: kernels.gaussian_points(kernels.sorted_uniform_points(n, 3), 0.05, 1e-2),
                                3000, 100, hmat_amd.HSS_SAMPLED, opts(32, 10)),
    # leaves of 15-16 rows, narrower than the ranks of their parents
    "small_leaves": (lambda n: cauchy(n), 1000, 20, hmat_amd.HSS_SAMPLED, opts(16, 10)),
    # leaves of 1000-1024 rows: four rows per thread in the leaf passes
    "wide_leaves": (lambda n: cauchy(n), 4000, 1024, hmat_amd.HSS_SAMPLED, opts(40, 10)),
    # nested bases from the N-wide leaf blocks, ranks chosen by tolerance
    "nested_tolerance": (lambda n: kernels.gaussian_points(kernels.sorted_uniform_points(n, 5), 0.1, 1.0),
                         2048, 64, True, opts(48, 10, tol=1e-10, fixed=False)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_one_launch_matvec(case):
    gen, n, leaf, mode, o = CASES[case]
    a = gen(n)
    h = hmat_amd.HSSMatrixOMP(opts=o, nested=mode)
    h.build_from_dense(a, leaf)
    G = h.to_dense()
    X = np.random.default_rng(5).standard_normal((n, 19))
    hmat_amd.profile_reset()
    hmat_amd.profile_enable(True)
    Y1 = np.column_stack([h.multiply(X[:, [j]]) for j in range(3)])
    hmat_amd.profile_enable(False)
    one = hmat_amd.profile_query("hss_matvec_one_launch")
    Y19 = h.multiply(X)  # groups of 16, 2 and 1 columns (8 and 4 where 16 does not fit)
    e1, e19, ec = rel(Y1, G @ X[:, :3]), rel(Y19, G @ X), rel(Y1, Y19[:, :3])
    print(f"\n{case}: root rank {h.root_rank()}, one-launch calls {one and one[0]}, one column vs dense {e1:.1e}, "
          f"19 columns vs dense {e19:.1e}, one vs grouped {ec:.1e}", flush=True)
    assert one is not None and one[0] == 3
    assert e1 < 1e-13
    assert e19 < 1e-13
    assert ec < 1e-14


def test_reference_ranks_take_the_sweeps():
    """Reference mode (identity transfers, ranks summed up the tree,
    hss_routines_omp.h:110-119) exceeds the one-launch kernel's rank limit
    of 64: single columns take the multi-launch sweeps and stay exact."""
    n = 2048
    a = cauchy(n)
    h = hmat_amd.HSSMatrixOMP(1e-12)
    h.build_from_dense(a, 64)
    assert h.root_rank() > 64
    G = h.to_dense()
    x = np.random.default_rng(6).standard_normal((n, 1))
    hmat_amd.profile_reset()
    hmat_amd.profile_enable(True)
    y = h.multiply(x)
    hmat_amd.profile_enable(False)
    assert hmat_amd.profile_query("hss_matvec_one_launch") is None
    assert rel(y, G @ x) < 1e-13


def test_wait_limit_reports_error():
    """A spin limit of one poll, set from the host, makes the persistent
    kernel's waits give up: every workgroup exits, the call returns
    HMAT_ERR_DEVICE, and the next call with the default limit gives the
    same product as before."""
    n = 16384
    h = hmat_amd.HSSMatrixOMP(opts=opts(32, 10), nested=hmat_amd.HSS_SAMPLED)
    h.build_from_dense(cauchy(n), 128)
    x = np.random.default_rng(7).standard_normal((n, 1))
    y0 = h.multiply(x)
    hmat_amd.set_fused_spin_limit(1)
    try:
        with pytest.raises(hmat_amd.HmatError) as e:
            h.multiply(x)
        assert e.value.status == 2  # HMAT_ERR_DEVICE
    finally:
        hmat_amd.set_fused_spin_limit(0)
    assert np.array_equal(h.multiply(x), y0)
