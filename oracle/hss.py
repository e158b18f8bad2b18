"""HSS restatement — oracle only.

Follows the serial ``HSSMatrix`` (include/hss_routines.h:10-307); the OMP
class (include/hss_routines_omp.h:11-375) computes the same algebra.

  * topology: root id 0; both children get ids before recursing
    (build_topo, hss_routines.h:293-306)
  * build (reverse id order, hss_routines.h:57-112):
      leaf:   D; U = compress_rsvd(A(I, :) with own columns zeroed).U;
              V = compress_rsvd(A(:, I)^T with own columns zeroed).U
      inner:  ranks add up, R = W = I, B12 = U_L^T A12 V_R placed at
              (0, rV_L), B21 = U_R^T A21 V_L placed at (rU_L, 0)
  * multiply: up-sweep g, down-sweep f = R f + B [g_l; g_r], y = D x + U f
    (hss_routines.h:118-192)
  * solve: x = z - Q B (I + K B)^-1 V^T z per node, Q = A_child^-1 U_child
    column by column (hss_routines.h:197-290)
"""
import numpy as np
import scipy.linalg as sla

from .lowrank import Options, compress


class HSSNode:
    __slots__ = ("id", "start", "end", "is_leaf", "left", "right", "parent",
                 "rank_U", "rank_V", "D", "U", "V", "R", "W", "B")

    def __init__(self, start=0, end=0):
        self.id = 0
        self.start, self.end = start, end
        self.is_leaf = False
        self.left = self.right = self.parent = None
        self.rank_U = self.rank_V = 0
        self.D = self.U = self.V = self.R = self.W = self.B = None


def _lu_solve(a, b):
    lu, piv = sla.lu_factor(a, check_finite=False)
    return sla.lu_solve((lu, piv), b, check_finite=False)


class HSSMatrix:
    def __init__(self, tol=1e-9, opts=None):
        self.tolerance = tol
        self.rsvd_rank = 64
        self.opts = opts
        self.nodes = []
        self.root = None
        self.N = 0

    def _options(self):
        if self.opts is not None:
            return self.opts
        return Options.reference(self.tolerance, self.rsvd_rank)

    def _build_topo(self, node, start, end, leaf):
        node.start, node.end = start, end
        if end - start <= leaf:
            node.is_leaf = True
            return
        mid = (start + end) // 2
        node.left = HSSNode()
        node.left.id = len(self.nodes)
        self.nodes.append(node.left)
        node.right = HSSNode()
        node.right.id = len(self.nodes)
        self.nodes.append(node.right)
        node.left.parent = node.right.parent = node
        self._build_topo(node.left, start, mid, leaf)
        self._build_topo(node.right, mid, end, leaf)

    def reconstruct_U(self, node):
        if node.is_leaf:
            return node.U
        ul, ur = self.reconstruct_U(node.left), self.reconstruct_U(node.right)
        big = np.zeros((ul.shape[0] + ur.shape[0], ul.shape[1] + ur.shape[1]))
        big[:ul.shape[0], :ul.shape[1]] = ul
        big[ul.shape[0]:, ul.shape[1]:] = ur
        return big @ node.R

    def reconstruct_V(self, node):
        if node.is_leaf:
            return node.V
        vl, vr = self.reconstruct_V(node.left), self.reconstruct_V(node.right)
        big = np.zeros((vl.shape[0] + vr.shape[0], vl.shape[1] + vr.shape[1]))
        big[:vl.shape[0], :vl.shape[1]] = vl
        big[vl.shape[0]:, vl.shape[1]:] = vr
        return big @ node.W

    # ---------------------------------------------------------------- build
    def build_from_dense(self, a, leaf_size, rng):
        n = a.shape[0]
        self.N = n
        self.nodes = []
        self.root = HSSNode()
        self.root.id = 0
        self.nodes.append(self.root)
        self._build_topo(self.root, 0, n, leaf_size)
        opts = self._options()
        for node in reversed(self.nodes):
            s, e = node.start, node.end
            if node.is_leaf:
                node.D = np.asfortranarray(a[s:e, s:e])
                row = np.array(a[s:e, :], order="F")
                row[:, s:e] = 0.0
                node.U = compress(row, opts, rng).U
                node.rank_U = node.U.shape[1]
                colt = np.array(a[:, s:e].T, order="F")
                colt[:, s:e] = 0.0
                node.V = compress(colt, opts, rng).U
                node.rank_V = node.V.shape[1]
                continue
            lc, rc = node.left, node.right
            node.rank_U = lc.rank_U + rc.rank_U
            node.rank_V = lc.rank_V + rc.rank_V
            node.R = np.eye(node.rank_U) if node.rank_U > 0 else np.zeros((0, 0))
            node.W = np.eye(node.rank_V) if node.rank_V > 0 else np.zeros((0, 0))
            if node.rank_U > 0 and node.rank_V > 0:
                mid = lc.end
                a12 = a[s:mid, mid:e]
                a21 = a[mid:e, s:mid]
                ul, vr = self.reconstruct_U(lc), self.reconstruct_V(rc)
                ur, vl = self.reconstruct_U(rc), self.reconstruct_V(lc)
                b12 = ul.T @ a12 @ vr
                b21 = ur.T @ a21 @ vl
                node.B = np.zeros((node.rank_U, node.rank_V))
                node.B[:b12.shape[0], lc.rank_V:lc.rank_V + b12.shape[1]] = b12
                node.B[lc.rank_U:lc.rank_U + b21.shape[0], :b21.shape[1]] = b21
            else:
                node.B = np.zeros((0, 0))

    # ------------------------------------------------------------- multiply
    def multiply(self, x):
        x = np.asarray(x, dtype=np.float64).reshape(self.N, -1)[:, :1]
        return self.multiply_multi(x)

    def multiply_multi(self, x):
        x = np.asarray(x, dtype=np.float64).reshape(self.N, -1)
        nrhs = x.shape[1]
        g = [None] * len(self.nodes)
        f = [None] * len(self.nodes)
        for i in range(len(self.nodes) - 1, -1, -1):
            node = self.nodes[i]
            if node.is_leaf:
                g[i] = node.V.T @ x[node.start:node.end] if node.rank_V > 0 \
                    else np.zeros((0, nrhs))
            elif node.rank_V == 0:
                g[i] = np.zeros((0, nrhs))
            else:
                g[i] = node.W.T @ np.vstack([g[node.left.id], g[node.right.id]])
        f[0] = np.zeros((self.root.rank_U, nrhs))
        for i, node in enumerate(self.nodes):
            if node.is_leaf:
                continue
            if node.rank_U == 0:
                f[node.left.id] = np.zeros((node.left.rank_U, nrhs))
                f[node.right.id] = np.zeros((node.right.rank_U, nrhs))
                continue
            gs = np.vstack([g[node.left.id], g[node.right.id]])
            tot = node.R @ f[i] + node.B @ gs
            rl = node.left.rank_U
            f[node.left.id] = tot[:rl]
            f[node.right.id] = tot[rl:rl + node.right.rank_U]
        y = np.zeros((self.N, nrhs))
        for i, node in enumerate(self.nodes):
            if node.is_leaf:
                yl = node.D @ x[node.start:node.end]
                if node.rank_U > 0:
                    yl = yl + node.U @ f[i]
                y[node.start:node.end] = yl
        return y

    # ---------------------------------------------------------------- solve
    def solve(self, b):
        b = np.asarray(b, dtype=np.float64).reshape(self.N, -1)
        return self._solve_multi(self.root, b)

    def _solve_multi(self, node, bm):
        out = np.zeros_like(bm)
        for j in range(bm.shape[1]):
            out[:, j:j + 1] = self._solve_rec(node, bm[:, j:j + 1])
        return out

    def _solve_rec(self, node, b):
        if node.is_leaf:
            return _lu_solve(node.D, b)
        len_l = node.left.end - node.left.start
        z1 = self._solve_rec(node.left, b[:len_l])
        z2 = self._solve_rec(node.right, b[len_l:])
        if node.rank_U == 0 or node.rank_V == 0:
            return np.vstack([z1, z2])
        uld, urd = self.reconstruct_U(node.left), self.reconstruct_U(node.right)
        q1 = self._solve_multi(node.left, uld)
        q2 = self._solve_multi(node.right, urd)
        vld, vrd = self.reconstruct_V(node.left), self.reconstruct_V(node.right)
        k = np.zeros((node.rank_V, node.rank_U))
        k[:node.left.rank_V, :node.left.rank_U] = vld.T @ q1
        k[node.left.rank_V:, node.left.rank_U:] = vrd.T @ q2
        m = k @ node.B + np.eye(node.rank_V)
        vtz = np.vstack([vld.T @ z1, vrd.T @ z2])
        gam = _lu_solve(m, vtz)
        bg = node.B @ gam
        rl = node.left.rank_U
        return np.vstack([z1 - q1 @ bg[:rl], z2 - q2 @ bg[rl:rl + node.right.rank_U]])

    # --------------------------------------------------- checking helpers
    def to_dense(self):
        a = np.zeros((self.N, self.N))
        for node in self.nodes:
            if node.is_leaf:
                a[node.start:node.end, node.start:node.end] = node.D
            elif node.B is not None and node.B.size:
                lc, rc = node.left, node.right
                ul, ur = self.reconstruct_U(lc), self.reconstruct_U(rc)
                vl, vr = self.reconstruct_V(lc), self.reconstruct_V(rc)
                b12 = node.B[:lc.rank_U, lc.rank_V:]
                b21 = node.B[lc.rank_U:, :lc.rank_V]
                a[lc.start:lc.end, rc.start:rc.end] = ul @ b12 @ vr.T
                a[rc.start:rc.end, lc.start:lc.end] = ur @ b21 @ vl.T
        return a

    def solve_dense(self, b):
        return np.linalg.solve(self.to_dense(), np.asarray(b).reshape(self.N, -1))


def omega_schedule(n, leaf, opts):
    """(n, k) of every Omega the serial HSS build draws, in draw order.

    Reverse id order over leaves; row block then column block, each
    len x N (hss_routines.h:57-77).
    """
    from .lowrank import uses_rsvd, sketch_width
    h = HSSMatrix()
    h.nodes = []
    h.root = HSSNode()
    h.nodes.append(h.root)
    h._build_topo(h.root, 0, n, leaf)
    out = []
    for node in reversed(h.nodes):
        if node.is_leaf:
            m_ = node.end - node.start
            for _ in range(2):
                if uses_rsvd(m_, n, opts):
                    out.append((n, sketch_width(m_, n, opts)))
    return out
