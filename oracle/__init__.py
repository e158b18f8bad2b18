"""CPU oracle for the HODLR / HSS hot path — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/scipy FP64, the algorithms of the reference
header-only library (chenxinye/hmat-omp):

  * ``rng``       — ``Matrix::Random`` (include/matrix.h:50-57): std::mt19937
                    + uniform_real_distribution<double>(0,1), column-major fill.
  * ``kernels``   — ``Kernels::generate_*`` (include/kernel.h:18-106) and the
                    point-based Gaussian / Cauchy kernels of the benchmark
                    configurations.
  * ``lowrank``   — ``Compression::compress_svd`` / ``compress_rsvd``
                    (include/compression.h:18-99), plus the fixed-rank,
                    power-iteration variant.
  * ``hodlr``     — ``HODLRMatrix`` (include/hodlr_routines.h:8-176), whose
                    algebra is that of ``HODLRMatrixOMP``
                    (include/hodlr_routines_omp.h:9-241).
  * ``hss``       — ``HSSMatrix`` (include/hss_routines.h:10-307), algebra of
                    ``HSSMatrixOMP`` (include/hss_routines_omp.h:11-375).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(the HIP library under ``hmat-omp_amd/``) never calls it.

Pinning (how each function is checked, so that it is not just re-typed):
  * rng: the C++ standard's known answer for mt19937 (10000th output of the
    default-seeded engine is 4123659995), numpy's independent MT19937, and
    golden vectors printed by libstdc++'s own uniform_real_distribution
    (tests/golden/).
  * kernels: closed forms at hand-computed entries.
  * compress_svd / compress_rsvd: numpy's gesdd SVD, orthonormality, the
    truncation rule on matrices with planted singular values, and the exact
    Eckart–Young error of the truncated SVD.
  * hodlr / hss build+multiply+solve: dense LAPACK (``numpy.linalg.solve``)
    on the dense reconstruction of the compressed operator, the Frobenius
    error against the true matrix, and the trivial tree (N <= leaf) which is
    a plain dense LU solve.
  * BLAS/LAPACK arithmetic of the reference (dgemm/dgesvd/dgetrf/dgetrs/
    dgetri, include/matrix.h:13-26) comes from an unpinned system library
    (CI: Ubuntu libblas/liblapack); scipy's LAPACK is used here and is
    itself checked against numpy's independent SVD/solve routines.
"""
