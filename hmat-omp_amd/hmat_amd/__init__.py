"""Python mirror of the reference classes over the C ABI (include/hmat_amd.h).

``HODLRMatrixOMP`` and ``HSSMatrixOMP`` keep the names, constructor
arguments and methods of the reference (include/hodlr_routines_omp.h:9-241,
include/hss_routines_omp.h:11-375): ``tol`` in the constructor, a public
``rsvd_rank`` (default 64), ``build_from_dense(A, leaf_size)``,
``multiply(x)`` and ``solve(b)``.  Matrices are numpy arrays (column-major
copies are made as needed).  Every call goes to the HIP library; there is no
CPU fallback: a missing library or device raises ``HmatError``.
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhmat_amd.so")

HMAT_OK = 0
STATUS_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_DEVICE", 3: "ERR_OOM", 4: "ERR_SINGULAR",
                5: "ERR_NONFINITE", 6: "ERR_STATE", 7: "ERR_NOT_CONVERGED"}

PREC_FP64, PREC_MIXED = 0, 1
OMEGA_UNIFORM, OMEGA_GAUSSIAN = 0, 1
KERNEL_GAUSSIAN, KERNEL_CAUCHY, KERNEL_REF_CAUCHY, KERNEL_REF_GAUSSIAN, KERNEL_REF_COULOMB = 1, 2, 3, 4, 5
KERNEL_GAUSSIAN_DSIGMA = 6
# HSS build modes (the `nested` argument of hmat_hss_build_*)
HSS_REFERENCE, HSS_NESTED, HSS_SAMPLED = 0, 1, 2
HSS_WOODBURY, HSS_ULV = 0, 1  # HSS solvers


class HmatError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


class Options(C.Structure):
    _fields_ = [("tol", C.c_double), ("rank", C.c_int32), ("fixed_rank", C.c_int32),
                ("oversample", C.c_int32), ("power_iters", C.c_int32),
                ("exact_below", C.c_int32), ("precision", C.c_int32),
                ("omega_dist", C.c_int32), ("seed", C.c_uint64)]


class Kernel(C.Structure):
    _fields_ = [("type", C.c_int32), ("p0", C.c_double), ("p1", C.c_double),
                ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double))]


_lib = None
_D = C.POINTER(C.c_double)


def _proto(lib):
    i64, i32, st = C.c_int64, C.c_int32, C.c_int
    vp = C.c_void_p
    sig = {
        "hmat_last_error": (C.c_char_p, []),
        "hmat_version": (C.c_char_p, []),
        "hmat_device_count": (i32, []),
        "hmat_options_reference": (None, [C.POINTER(Options), C.c_double]),
        "hmat_profile_enable": (None, [i32]),
        "hmat_profile_reset": (None, []),
        "hmat_autotune_enable": (None, [i32]),
        "hmat_autotune_entries": (i64, []),
        "hmat_profile_query": (i32, [C.c_char_p, C.POINTER(i64), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
        "hmat_hodlr_create": (st, [C.POINTER(vp), C.POINTER(Options)]),
        "hmat_hodlr_destroy": (st, [vp]),
        "hmat_hodlr_build_dense": (st, [vp, _D, i64, i64, i32, _D, i64]),
        "hmat_hodlr_build_dense_device": (st, [vp, vp, i64, i64, i32, _D, i64]),
        "hmat_hodlr_build_kernel": (st, [vp, C.POINTER(Kernel), i64, i32]),
        "hmat_hodlr_build_kernel_device": (st, [vp, C.POINTER(Kernel), i64, i32]),
        "hmat_kernel_matvec_device": (st, [C.POINTER(Kernel), i64, vp, i64, i32, vp, i64]),
        "hmat_hodlr_omega_length": (i64, [C.POINTER(Options), i64, i32]),
        "hmat_hodlr_build_kernel_batch_device": (st, [vp, C.POINTER(Kernel), C.POINTER(i64), i32, i32]),
        "hmat_hodlr_batch_note": (C.c_char_p, [vp]),
        "hmat_hss_matvec_device": (st, [vp, vp, i64, i32, vp, i64]),
        "hmat_hss_solve_device": (st, [vp, vp, i64, i32, vp, i64]),
        "hmat_hodlr_batch_solve_device": (st, [vp, C.POINTER(Kernel), C.POINTER(i64), i32, i32, vp, i64, i32,
                                               vp, i64, i32, C.POINTER(i32)]),
        "hmat_hodlr_matvec": (st, [vp, _D, i64, i32, _D, i64]),
        "hmat_hodlr_matvec_device": (st, [vp, vp, i64, i32, vp, i64]),
        "hmat_hodlr_matvec_device_noalloc": (st, [vp, vp, i64, i32, vp, i64]),
        "hmat_hodlr_matvec_noalloc_capacity": (i64, []),
        "hmat_hodlr_factor": (st, [vp]),
        "hmat_hodlr_solve": (st, [vp, _D, i64, i32, _D, i64]),
        "hmat_hodlr_solve_device": (st, [vp, vp, i64, i32, vp, i64]),
        "hmat_hodlr_solve_transpose": (st, [vp, _D, i64, i32, _D, i64]),
        "hmat_hodlr_solve_transpose_device": (st, [vp, vp, i64, i32, vp, i64]),
        "hmat_hodlr_matvec_transpose": (st, [vp, _D, i64, i32, _D, i64]),
        "hmat_hodlr_set_refinement": (st, [vp, C.c_double, i32]),
        "hmat_hodlr_last_refinement": (st, [vp, C.POINTER(i32), C.POINTER(C.c_double)]),
        "hmat_hodlr_size": (i64, [vp]),
        "hmat_hodlr_levels": (i32, [vp]),
        "hmat_hodlr_level_ranks": (st, [vp, i32, C.POINTER(i32), i32, C.POINTER(i32)]),
        "hmat_hodlr_to_dense": (st, [vp, _D, i64]),
        "hmat_hodlr_set_stream": (st, [vp, vp]),
        "hmat_hodlr_save": (st, [vp, C.c_char_p]),
        "hmat_hodlr_set_solve_graph": (st, [vp, i32]),
        "hmat_hodlr_set_fused_solve": (st, [vp, i32]),
        "hmat_set_fused_spin_limit": (None, [i64]),
        "hmat_hodlr_set_group_bytes": (st, [vp, C.c_double]),
        "hmat_hodlr_load": (st, [vp, C.c_char_p]),
        "hmat_hss_create": (st, [C.POINTER(vp), C.POINTER(Options)]),
        "hmat_hss_destroy": (st, [vp]),
        "hmat_hss_build_dense": (st, [vp, _D, i64, i64, i32, i32, _D, i64]),
        "hmat_hss_build_kernel": (st, [vp, C.POINTER(Kernel), i64, i32, i32]),
        "hmat_hss_omega_length": (i64, [C.POINTER(Options), i64, i32]),
        "hmat_hss_sampled_omega_length": (i64, [C.POINTER(Options), i64]),
        "hmat_hss_build_kernel_device": (st, [vp, C.POINTER(Kernel), i64, i32, i32]),
        "hmat_hss_matvec": (st, [vp, _D, i64, i32, _D, i64]),
        "hmat_hss_factor": (st, [vp]),
        "hmat_hss_solve": (st, [vp, _D, i64, i32, _D, i64]),
        "hmat_hss_size": (i64, [vp]),
        "hmat_hss_root_rank": (i32, [vp]),
        "hmat_hss_to_dense": (st, [vp, _D, i64]),
        "hmat_hss_set_refinement": (st, [vp, C.c_double, i32]),
        "hmat_hss_set_solver": (st, [vp, i32]),
        "hmat_hss_last_refinement": (st, [vp, C.POINTER(i32), C.POINTER(C.c_double)]),
        "hmat_hss_save": (st, [vp, C.c_char_p]),
        "hmat_hss_load": (st, [vp, C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load(path=None):
    """Load (once) and return the HIP library; raises if it is missing."""
    global _lib
    if _lib is None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise HmatError(2, f"{p} not built (run hmat-omp_amd/build.py)")
        _one_hip_runtime()
        _lib = _proto(C.CDLL(p))
    return _lib


def _one_hip_runtime():
    """One HIP runtime per process.  torch's ROCm wheel carries its own
    libamdhip64 (torch/lib, soname libamdhip64.so.7) and its libc10_hip asks
    for the unversioned name; the engine links the system libamdhip64.so.7.
    Loaded after torch, the engine binds to torch's copy (same soname); loaded
    first, it makes torch open a second runtime later, and that one finds no
    device.  So when torch is installed it is loaded first; device pointers
    from torch tensors and from the engine then live in one runtime."""
    if "torch" in sys.modules:
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def _check(st):
    if st != HMAT_OK:
        raise HmatError(st, load().hmat_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(_D) if a is not None else None


def _fortran(a, rows=None):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return np.asfortranarray(a)


def reference_options(tol):
    o = Options()
    load().hmat_options_reference(C.byref(o), tol)
    return o


def make_options(tol=1e-6, rank=64, fixed_rank=False, oversample=10, power_iters=0,
                 exact_below=256, precision=PREC_FP64, omega_dist=OMEGA_UNIFORM, seed=0x5EED):
    return Options(tol, rank, int(fixed_rank), oversample, power_iters, exact_below, precision,
                   omega_dist, seed)


def make_kernel_device(ktype, x_ptr, y_ptr=None, p0=0.0, p1=0.0):
    """Kernel descriptor whose points are device pointers (ints)."""
    return Kernel(ktype, p0, p1, C.cast(C.c_void_p(x_ptr), _D),
                  C.cast(C.c_void_p(y_ptr), _D) if y_ptr else None)


def kernel_matvec_device(kernel, n, x_ptr, ldx, nrhs, y_ptr, ldy):
    _check(load().hmat_kernel_matvec_device(C.byref(kernel), n, x_ptr, ldx, nrhs, y_ptr, ldy))


def make_kernel(ktype, x, y=None, p0=0.0, p1=0.0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    k = Kernel(ktype, p0, p1, _ptr(x), _ptr(y))
    k._keep = (x, y)
    return k


def autotune_enable(on=True):
    """Time candidate tilings on the first launch of each shape class."""
    load().hmat_autotune_enable(int(bool(on)))


def autotune_entries():
    return load().hmat_autotune_entries()


def set_fused_spin_limit(n):
    """Spin limit of the persistent kernels' waits: the HODLR apply-inverse
    and the HSS matvec (0: default)."""
    load().hmat_set_fused_spin_limit(int(n))


def profile_enable(on=True):
    load().hmat_profile_enable(int(on))


def profile_reset():
    load().hmat_profile_reset()


def profile_query(name):
    """(launches, total ms, total algorithmic work) or None."""
    n, ms, w = C.c_int64(), C.c_double(), C.c_double()
    if not load().hmat_profile_query(name.encode(), C.byref(n), C.byref(ms), C.byref(w)):
        return None
    return n.value, ms.value, w.value


class _Base:
    _prefix = None

    def __init__(self, tol, opts):
        lib = load()
        self._lib = lib
        if opts is None:
            opts = reference_options(tol)
        self.opts = opts
        self.tolerance = opts.tol
        self.rsvd_rank = opts.rank
        self._h = C.c_void_p()
        self.N = 0

    def _sync_opts(self):
        # the reference exposes rsvd_rank as a public member set after construction
        if self.rsvd_rank != self.opts.rank:
            self.opts.rank = self.rsvd_rank
            self._recreate()

    def _f(self, name):
        return getattr(self._lib, f"hmat_{self._prefix}_{name}")

    def _recreate(self):
        if self._h:
            _check(self._f("destroy")(self._h))
        self._h = C.c_void_p()
        _check(self._f("create")(C.byref(self._h), C.byref(self.opts)))

    def __del__(self):
        try:
            if self._h:
                self._f("destroy")(self._h)
        except Exception:
            pass

    def multiply(self, x):
        """Y = A_H x for every column of x."""
        x = _fortran(x)
        y = np.zeros((self.N, x.shape[1]), order="F")
        _check(self._f("matvec")(self._h, _ptr(x), x.shape[0], x.shape[1], _ptr(y), self.N))
        return y

    def factor(self):
        _check(self._f("factor")(self._h))

    def solve(self, b):
        b = _fortran(b)
        x = np.zeros((self.N, b.shape[1]), order="F")
        _check(self._f("solve")(self._h, _ptr(b), b.shape[0], b.shape[1], _ptr(x), self.N))
        return x

    def to_dense(self):
        a = np.zeros((self.N, self.N), order="F")
        _check(self._f("to_dense")(self._h, _ptr(a), self.N))
        return a


class HODLRMatrixOMP(_Base):
    """hodlr_routines_omp.h:9-241 on the GPU."""
    _prefix = "hodlr"

    def __init__(self, tol=1e-6, opts=None):
        super().__init__(tol, opts)
        self._recreate()

    def build_from_dense(self, a, leaf_size, omega=None):
        self._sync_opts()
        a = _fortran(a)
        n = a.shape[0]
        om = None if omega is None else np.ascontiguousarray(omega, dtype=np.float64)
        olen = 0 if om is None else om.size
        _check(self._lib.hmat_hodlr_build_dense(self._h, _ptr(a), n, n, leaf_size, _ptr(om), olen))
        self.N = n

    def build_kernel(self, kernel, n, leaf_size):
        self._sync_opts()
        _check(self._lib.hmat_hodlr_build_kernel(self._h, C.byref(kernel), n, leaf_size))
        self.N = n

    def build_kernel_device(self, kernel, n, leaf_size):
        self._sync_opts()
        _check(self._lib.hmat_hodlr_build_kernel_device(self._h, C.byref(kernel), n, leaf_size))
        self.N = n

    def solve_device(self, b_ptr, ldb, nrhs, x_ptr, ldx):
        _check(self._lib.hmat_hodlr_solve_device(self._h, b_ptr, ldb, nrhs, x_ptr, ldx))

    def set_solve_graph(self, on):
        """Replay solves from a captured HIP graph (default) or launch eagerly."""
        _check(self._lib.hmat_hodlr_set_solve_graph(self._h, int(bool(on))))

    def set_fused_solve(self, mode):
        """Apply-inverse as one persistent launch per four columns: 1 (or True,
        the default) where that is the faster path, 0 (False) never, 2 whenever
        the device holds its grid."""
        _check(self._lib.hmat_hodlr_set_fused_solve(self._h, int(mode)))

    def set_group_bytes(self, nbytes):
        """Scratch budget of one group of depths built together (0: default)."""
        _check(self._lib.hmat_hodlr_set_group_bytes(self._h, float(nbytes)))

    def save(self, path):
        """Write the built operator to `path` (see hmat_hodlr_save)."""
        _check(self._lib.hmat_hodlr_save(self._h, str(path).encode()))

    def load(self, path):
        """Replace this handle's operator and options with the file's."""
        _check(self._lib.hmat_hodlr_load(self._h, str(path).encode()))
        self.N = self._lib.hmat_hodlr_size(self._h)

    def build_kernel_batch_device(self, kernel, sizes, leaf_size):
        """Independent systems of the given sizes on consecutive points (a
        block-diagonal operator, built and solved for all systems at once)."""
        self._sync_opts()
        sz = (C.c_int64 * len(sizes))(*[int(v) for v in sizes])
        _check(self._lib.hmat_hodlr_build_kernel_batch_device(self._h, C.byref(kernel), sz, len(sizes),
                                                              leaf_size))
        self.N = int(sum(int(v) for v in sizes))

    BATCH_PATHS = {0: "eager", 1: "recorded", 2: "replayed", 3: "redone"}

    def batch_solve_device(self, kernel, sizes, leaf_size, b_ptr, ldb, ncols, x_ptr, ldx, graph=True):
        """Build + factor + solve of a batch in one call (device arrays);
        with graph=True the work is recorded into one HIP graph on the first
        call for a shape and replayed on later ones.  Returns how the call
        ran: "eager", "recorded", "replayed" or "redone"."""
        self._sync_opts()
        sz = (C.c_int64 * len(sizes))(*[int(v) for v in sizes])
        path = C.c_int32(-1)
        _check(self._lib.hmat_hodlr_batch_solve_device(self._h, C.byref(kernel), sz, len(sizes), leaf_size,
                                                       b_ptr, ldb, ncols, x_ptr, ldx, int(bool(graph)),
                                                       C.byref(path)))
        self.N = int(sum(int(v) for v in sizes))
        return self.BATCH_PATHS[path.value]

    def batch_note(self):
        """Why the last graph batch call neither recorded nor replayed ("" if it did)."""
        return self._lib.hmat_hodlr_batch_note(self._h).decode()

    def multiply_device(self, x_ptr, ldx, nrhs, y_ptr, ldy):
        _check(self._lib.hmat_hodlr_matvec_device(self._h, x_ptr, ldx, nrhs, y_ptr, ldy))

    def multiply_device_noalloc(self, x_ptr, ldx, nrhs, y_ptr, ldy):
        """The same product through the allocation-free entry point."""
        _check(self._lib.hmat_hodlr_matvec_device_noalloc(self._h, x_ptr, ldx, nrhs, y_ptr, ldy))

    def solve_transpose(self, b):
        """X = A_H^{-T} B (adjoint solve; no reference counterpart)."""
        b = _fortran(b)
        x = np.zeros((self.N, b.shape[1]), order="F")
        _check(self._lib.hmat_hodlr_solve_transpose(self._h, _ptr(b), b.shape[0], b.shape[1], _ptr(x), self.N))
        return x

    def solve_transpose_device(self, b_ptr, ldb, nrhs, x_ptr, ldx):
        _check(self._lib.hmat_hodlr_solve_transpose_device(self._h, b_ptr, ldb, nrhs, x_ptr, ldx))

    def multiply_transpose(self, x):
        """Y = A_H^T X."""
        x = _fortran(x)
        y = np.zeros((self.N, x.shape[1]), order="F")
        _check(self._lib.hmat_hodlr_matvec_transpose(self._h, _ptr(x), x.shape[0], x.shape[1], _ptr(y), self.N))
        return y

    def omega_length(self, n, leaf_size):
        self._sync_opts()
        return self._lib.hmat_hodlr_omega_length(C.byref(self.opts), n, leaf_size)

    def levels(self):
        return self._lib.hmat_hodlr_levels(self._h)

    def level_ranks(self, lvl):
        cnt = C.c_int32()
        _check(self._lib.hmat_hodlr_level_ranks(self._h, lvl, None, 0, C.byref(cnt)))
        out = (C.c_int32 * cnt.value)()
        _check(self._lib.hmat_hodlr_level_ranks(self._h, lvl, out, cnt.value, C.byref(cnt)))
        return list(out)

    def set_refinement(self, rtol, max_iters=50):
        _check(self._lib.hmat_hodlr_set_refinement(self._h, rtol, max_iters))

    def last_refinement(self):
        it, res = C.c_int32(), C.c_double()
        _check(self._lib.hmat_hodlr_last_refinement(self._h, C.byref(it), C.byref(res)))
        return it.value, res.value


class HSSMatrixOMP(_Base):
    """hss_routines_omp.h:11-375 on the GPU."""
    _prefix = "hss"

    def __init__(self, tol=1e-9, opts=None, nested=False):
        """nested: False / HSS_REFERENCE (identity transfers, the reference
        operator), True / HSS_NESTED (compressed transfers) or HSS_SAMPLED
        (nested bases from one sample pass per side)."""
        if opts is None:
            opts = reference_options(tol)
        super().__init__(tol, opts)
        self.nested = nested
        self._recreate()

    def build_from_dense(self, a, leaf_size, omega=None):
        self._sync_opts()
        a = _fortran(a)
        n = a.shape[0]
        om = None if omega is None else np.ascontiguousarray(omega, dtype=np.float64)
        olen = 0 if om is None else om.size
        _check(self._lib.hmat_hss_build_dense(self._h, _ptr(a), n, n, leaf_size, int(self.nested),
                                              _ptr(om), olen))
        self.N = n

    def build_kernel(self, kernel, n, leaf_size):
        self._sync_opts()
        _check(self._lib.hmat_hss_build_kernel(self._h, C.byref(kernel), n, leaf_size,
                                               int(self.nested)))
        self.N = n

    def build_kernel_device(self, kernel, n, leaf_size):
        """build_kernel with the kernel's points in device memory."""
        self._sync_opts()
        _check(self._lib.hmat_hss_build_kernel_device(self._h, C.byref(kernel), n, leaf_size,
                                                      int(self.nested)))
        self.N = n

    def multiply_device(self, x_ptr, ldx, nrhs, y_ptr, ldy):
        _check(self._lib.hmat_hss_matvec_device(self._h, x_ptr, ldx, nrhs, y_ptr, ldy))

    def solve_device(self, b_ptr, ldb, nrhs, x_ptr, ldx):
        _check(self._lib.hmat_hss_solve_device(self._h, b_ptr, ldb, nrhs, x_ptr, ldx))

    def omega_length(self, n, leaf_size):
        self._sync_opts()
        if int(self.nested) == HSS_SAMPLED:
            return self._lib.hmat_hss_sampled_omega_length(C.byref(self.opts), n)
        return self._lib.hmat_hss_omega_length(C.byref(self.opts), n, leaf_size)

    def root_rank(self):
        return self._lib.hmat_hss_root_rank(self._h)

    def set_refinement(self, rtol, max_iters=30):
        """FP64 refinement of solves to max relative residual rtol (0: off)."""
        _check(self._lib.hmat_hss_set_refinement(self._h, rtol, max_iters))

    def last_refinement(self):
        it, res = C.c_int32(), C.c_double()
        _check(self._lib.hmat_hss_last_refinement(self._h, C.byref(it), C.byref(res)))
        return it.value, res.value

    def set_solver(self, solver):
        """HSS_WOODBURY (the reference's push-through solve, default) or
        HSS_ULV (orthogonal ULV factorisation, backward stable); drops the
        factorisation."""
        _check(self._lib.hmat_hss_set_solver(self._h, solver))

    def save(self, path):
        """Write the built operator to `path` (see hmat_hss_save)."""
        _check(self._lib.hmat_hss_save(self._h, str(path).encode()))

    def load(self, path):
        """Replace this handle's operator and options with the file's."""
        _check(self._lib.hmat_hss_load(self._h, str(path).encode()))
        self.N = self._lib.hmat_hss_size(self._h)
