"""HODLR restatement — oracle only.

Follows the serial ``HODLRMatrix`` (include/hodlr_routines.h:8-176), whose
Omega draw order is deterministic (post-order; A12 then A21,
hodlr_routines.h:43-53).  ``HODLRMatrixOMP`` (include/hodlr_routines_omp.h)
computes the same algebra with OpenMP tasks:

  * build:    bisection mid=(s+e)/2, leaf if len <= leaf   (omp.h:33-79)
  * multiply: y_l += U_L B12 V_R^T x_r, y_r += U_R B21 V_L^T x_l (omp.h:91-135)
  * solve:    Woodbury with C = [[B12^-1, V_R^T Q2], [V_L^T Q1, B21^-1]],
              Q_i = A_ii^-1 U_i computed recursively           (omp.h:147-240)

Extras for checking the engine (not in the reference): multi-column
multiply, dense reconstruction of the compressed operator, and a dense LU
solve of that reconstruction.
"""
import numpy as np
import scipy.linalg as sla

from .lowrank import Options, compress


class Node:
    __slots__ = ("start", "end", "is_leaf", "left", "right", "D", "U", "V",
                 "B12", "B21")

    def __init__(self, start, end):
        self.start, self.end = start, end
        self.is_leaf = False
        self.left = self.right = None
        self.D = self.U = self.V = self.B12 = self.B21 = None


def _lu_solve(a, b):
    """Matrix::solve_lu (matrix.h:120-131): dgetrf + dgetrs."""
    lu, piv = sla.lu_factor(a, check_finite=False)
    return sla.lu_solve((lu, piv), b, check_finite=False)


def _inverse(a):
    """Matrix::inverse (matrix.h:133-149): dgetrf + dgetri."""
    return sla.inv(a, check_finite=False)


class HODLRMatrix:
    def __init__(self, tol=1e-6, opts=None):
        self.tolerance = tol
        self.rsvd_rank = 64
        self.opts = opts
        self.root = None
        self.N = 0

    def _options(self):
        if self.opts is not None:
            return self.opts
        return Options.reference(self.tolerance, self.rsvd_rank)

    # ---------------------------------------------------------------- build
    def build_from_dense(self, a, leaf_size, rng):
        """hodlr_routines.h:21-64; ``rng`` supplies Matrix::Random."""
        self.N = a.shape[0]
        self.leaf_size = leaf_size
        self.root = Node(0, self.N)
        opts = self._options()
        self._build(self.root, a, leaf_size, opts, rng)

    def _build(self, node, a, leaf, opts, rng):
        s, e = node.start, node.end
        if e - s <= leaf:
            node.is_leaf = True
            node.D = np.asfortranarray(a[s:e, s:e])
            return
        mid = (s + e) // 2
        node.left, node.right = Node(s, mid), Node(mid, e)
        self._build(node.left, a, leaf, opts, rng)
        self._build(node.right, a, leaf, opts, rng)
        lr12 = compress(np.asfortranarray(a[s:mid, mid:e]), opts, rng)
        lr21 = compress(np.asfortranarray(a[mid:e, s:mid]), opts, rng)
        node.left.U, node.right.V, node.B12 = lr12.U, lr12.V, lr12.B
        node.right.U, node.left.V, node.B21 = lr21.U, lr21.V, lr21.B

    # ------------------------------------------------------------- multiply
    def multiply(self, x):
        """Single column, as the reference (reads column 0, omp.h:96,115)."""
        x = np.asarray(x, dtype=np.float64).reshape(self.N, -1)
        return self._mult(self.root, x[:, :1])

    def multiply_multi(self, x):
        x = np.asarray(x, dtype=np.float64).reshape(self.N, -1)
        return self._mult(self.root, x)

    def _mult(self, node, x):
        s, e = node.start, node.end
        if node.is_leaf:
            return node.D @ x[s:e]
        mid = node.left.end
        y_l = self._mult(node.left, x)
        y_r = self._mult(node.right, x)
        y_l = y_l + node.left.U @ (node.B12 @ (node.right.V.T @ x[mid:e]))
        y_r = y_r + node.right.U @ (node.B21 @ (node.left.V.T @ x[s:mid]))
        return np.vstack([y_l, y_r])

    # ---------------------------------------------------------------- solve
    def solve(self, b):
        """hodlr_routines.h:137-176 (recursive Woodbury, multi-column)."""
        b = np.asarray(b, dtype=np.float64).reshape(self.N, -1)
        return self._solve(self.root, b)

    def _solve(self, node, b):
        if node.is_leaf:
            return _lu_solve(node.D, b)
        len_l = node.left.end - node.left.start
        z1 = self._solve(node.left, b[:len_l])
        z2 = self._solve(node.right, b[len_l:])
        vt_z_1 = node.right.V.T @ z2
        vt_z_2 = node.left.V.T @ z1
        q1 = self._solve(node.left, node.left.U)
        q2 = self._solve(node.right, node.right.U)
        r1, r2 = node.B12.shape[0], node.B21.shape[0]
        c = np.zeros((r1 + r2, r1 + r2))
        c[:r1, :r1] = _inverse(node.B12)
        c[r1:, r1:] = _inverse(node.B21)
        c[:r1, r1:] = node.right.V.T @ q2
        c[r1:, :r1] = node.left.V.T @ q1
        rhs = np.vstack([vt_z_1, vt_z_2])
        gamma = _lu_solve(c, rhs)
        return np.vstack([z1 - q1 @ gamma[:r1], z2 - q2 @ gamma[r1:]])

    # --------------------------------------------------- checking helpers
    def to_dense(self):
        a = np.zeros((self.N, self.N))
        self._dense(self.root, a)
        return a

    def _dense(self, node, a):
        s, e = node.start, node.end
        if node.is_leaf:
            a[s:e, s:e] = node.D
            return
        mid = node.left.end
        a[s:mid, mid:e] = node.left.U @ node.B12 @ node.right.V.T
        a[mid:e, s:mid] = node.right.U @ node.B21 @ node.left.V.T
        self._dense(node.left, a)
        self._dense(node.right, a)

    def solve_dense(self, b):
        return np.linalg.solve(self.to_dense(), np.asarray(b).reshape(self.N, -1))

    def nodes_postorder(self):
        out = []

        def walk(n):
            if not n.is_leaf:
                walk(n.left)
                walk(n.right)
            out.append(n)
        walk(self.root)
        return out

    def ranks_by_level(self):
        """{level: [(rank12, rank21), ...]} for internal nodes."""
        out = {}

        def walk(n, lvl):
            if n.is_leaf:
                return
            out.setdefault(lvl, []).append((n.B12.shape[0], n.B21.shape[0]))
            walk(n.left, lvl + 1)
            walk(n.right, lvl + 1)
        walk(self.root, 0)
        return out


def omega_schedule(n, leaf, opts):
    """(n, k) of every Omega the serial build draws, in draw order.

    Post-order over internal nodes; A12 (m=len_l, n=len_r) then A21
    (m=len_r, n=len_l); only blocks that take the randomized branch.
    """
    from .lowrank import uses_rsvd, sketch_width
    out = []

    def walk(s, e):
        if e - s <= leaf:
            return
        mid = (s + e) // 2
        walk(s, mid)
        walk(mid, e)
        for m_, n_ in ((mid - s, e - mid), (e - mid, mid - s)):
            if uses_rsvd(m_, n_, opts):
                out.append((n_, sketch_width(m_, n_, opts)))
    walk(0, n)
    return out


def rel_error(ref, res):
    """calc_rel_error (src/main_hodlr.cpp:26-35): Frobenius relative error."""
    ref = np.asarray(ref).ravel()
    res = np.asarray(res).ravel()
    den = np.dot(ref, ref)
    if den == 0:
        return 0.0
    d = ref - res
    return float(np.sqrt(np.dot(d, d)) / np.sqrt(den))
