"""Randomized nested HSS construction — oracle only (test infrastructure).

No reference counterpart: the reference builds HSS leaf bases from the whole
N-wide off-diagonal row / column blocks with identity transfers
(include/hss_routines_omp.h:69-95, 110-119), which needs O(N^2 k / leaf)
memory.  This restates the single-pass randomized construction of
P.-G. Martinsson, "A fast randomized algorithm for computing a hierarchically
semiseparable representation of a matrix", SIMAX 32(4), 2011, in its
orthonormal-basis form, which the engine's sampled HSS build follows:

  samples    Y = A Omega, Z = A^T Psi (N x s, Gaussian)
  leaf i     Yh_i = A(I, I^c) Omega(I^c), Zh_i = A(I^c, I)^T Psi(I^c);
             U_i = leading left singular vectors of Yh_i (rank rule below),
             yh_i = U_i^T Yh_i, om_i = V_i^T Omega(I), ps_i = U_i^T Psi(I)
             (V_i, zh_i likewise from Zh_i)
  siblings   B_ab = U_a^T A(I_a, I_b) V_b with the children's full bases
  parent p   M_p = [yh_a - B_ab om_b; yh_b - B_ba om_a] -> R_p, yh_p = R_p^T M_p
             N_p = [zh_a - B_ba^T ps_b; zh_b - B_ab^T ps_a] -> W_p, zh_p
             om_p = W_p^T [om_a; om_b], ps_p = R_p^T [ps_a; ps_b]
             full bases U_p = diag(U_a, U_b) R_p, V_p = diag(V_a, V_b) W_p

Rank rule (the engine's, from compression.h:42-46 applied to the sample's
singular values S): fixed rank r -> min(r, len(S)); tolerance -> #{S_j >
tol S_0}, at least 1, and 1 when S_0 <= 1e-15.  The tree is the
reference's bisection (mid = (s + e) / 2, leaf when len <= leaf).
"""
import numpy as np


class Node:
    __slots__ = ("start", "len", "depth", "left", "right")

    def __init__(self, start, length, depth):
        self.start, self.len, self.depth = start, length, depth
        self.left = self.right = -1


def bisect(n, leaf):
    nodes = []

    def rec(s, e, d):
        i = len(nodes)
        nodes.append(Node(s, e - s, d))
        if e - s > leaf:
            m = (s + e) // 2
            l_ = rec(s, m, d + 1)
            r_ = rec(m, e, d + 1)
            nodes[i].left, nodes[i].right = l_, r_
        return i

    rec(0, n, 0)
    return nodes


def sample_rank(s, fixed_rank, tol):
    if fixed_rank:
        return min(fixed_rank, len(s))
    if not s[0] > 1e-15:
        return 1
    return max(1, int(np.sum(s > tol * s[0])))


def _basis(m, fixed_rank, tol):
    u, s, vt = np.linalg.svd(m, full_matrices=False)
    r = sample_rank(s, fixed_rank, tol)
    return u[:, :r], s[:r, None] * vt[:r]


class HSSSampled:
    """Nested HSS generators of a dense matrix from one sample pass per side."""

    def __init__(self, rank=0, tol=1e-10):
        self.fixed_rank, self.tol = rank, tol

    def build_from_dense(self, a, leaf, omega, psi, subtract_siblings=True):
        """subtract_siblings=False leaves out the sibling terms of the parent
        samples (B_ab om_b, ...): a planted mistake for the tests only."""
        n = a.shape[0]
        self.N = n
        self.nodes = nodes = bisect(n, leaf)
        nn = len(nodes)
        y, z = a @ omega, a.T @ psi
        self.D = [None] * nn
        self.U = [None] * nn   # full (expanded) bases
        self.V = [None] * nn
        self.R = [None] * nn
        self.W = [None] * nn
        self.B = [None] * nn   # (B_ab, B_ba) at the parent
        yh, zh, om, ps = [None] * nn, [None] * nn, [None] * nn, [None] * nn
        fr, tol = self.fixed_rank, self.tol
        dmax = max(nd.depth for nd in nodes)
        for d in range(dmax, -1, -1):
            for i, nd in enumerate(nodes):
                if nd.depth != d:
                    continue
                rows = slice(nd.start, nd.start + nd.len)
                if nd.left < 0:
                    self.D[i] = a[rows, rows].copy()
                    if d == 0:
                        continue
                    yl = y[rows] - self.D[i] @ omega[rows]
                    zl = z[rows] - self.D[i].T @ psi[rows]
                    self.U[i], yh[i] = _basis(yl, fr, tol)
                    self.V[i], zh[i] = _basis(zl, fr, tol)
                    om[i] = self.V[i].T @ omega[rows]
                    ps[i] = self.U[i].T @ psi[rows]
                    continue
                ia, ib = nd.left, nd.right
                na, nb = nodes[ia], nodes[ib]
                ra = slice(na.start, na.start + na.len)
                rb = slice(nb.start, nb.start + nb.len)
                bab = self.U[ia].T @ a[ra, rb] @ self.V[ib]
                bba = self.U[ib].T @ a[rb, ra] @ self.V[ia]
                self.B[i] = (bab, bba)
                if d == 0:
                    continue
                c = 1.0 if subtract_siblings else 0.0
                m = np.vstack([yh[ia] - c * (bab @ om[ib]), yh[ib] - c * (bba @ om[ia])])
                q = np.vstack([zh[ia] - c * (bba.T @ ps[ib]), zh[ib] - c * (bab.T @ ps[ia])])
                self.R[i], yh[i] = _basis(m, fr, tol)
                self.W[i], zh[i] = _basis(q, fr, tol)
                om[i] = self.W[i].T @ np.vstack([om[ia], om[ib]])
                ps[i] = self.R[i].T @ np.vstack([ps[ia], ps[ib]])
                ua = self.U[ia].shape[1]
                va = self.V[ia].shape[1]
                self.U[i] = np.vstack([self.U[ia] @ self.R[i][:ua], self.U[ib] @ self.R[i][ua:]])
                self.V[i] = np.vstack([self.V[ia] @ self.W[i][:va], self.V[ib] @ self.W[i][va:]])
        return self

    def to_dense(self):
        a = np.zeros((self.N, self.N))
        for i, nd in enumerate(self.nodes):
            if nd.left < 0:
                a[nd.start:nd.start + nd.len, nd.start:nd.start + nd.len] = self.D[i]
                continue
            na, nb = self.nodes[nd.left], self.nodes[nd.right]
            bab, bba = self.B[i]
            a[na.start:na.start + na.len, nb.start:nb.start + nb.len] = self.U[nd.left] @ bab @ self.V[nd.right].T
            a[nb.start:nb.start + nb.len, na.start:na.start + na.len] = self.U[nd.right] @ bba @ self.V[nd.left].T
        return a

    def ranks_by_depth(self):
        """{depth: [rank_U of each non-root node at that depth]}"""
        out = {}
        for i, nd in enumerate(self.nodes):
            if nd.depth > 0:
                out.setdefault(nd.depth, []).append(self.U[i].shape[1])
        return out
