"""ULV factorisation and solve of nested HSS generators — oracle only (test
infrastructure).

No reference counterpart: the reference solves HSS systems with the
push-through Woodbury formula (include/hss_routines_omp.h:247-346).  This
restates the ULV elimination of S. Chandrasekaran, M. Gu and T. Pals, "A
fast ULV decomposition solver for hierarchically semiseparable
representations", SIMAX 28(3), 2006, in the form the engine runs, for
generators whose row bases and transfers have orthonormal columns (all of
the engine's builds give such bases).

Factor, children before parents.  Node i has an n x n block D_i, a row basis
U_i (n x r) and a column basis V_i (n x s); the root has no bases (r = s = 0).
  Householder QR (LAPACK convention) U_i = H_U [R_u; 0], S = sign(diag R_u),
  Qt = diag(S, I) H_U^T  (so Qt U_i = [S R_u; 0], S R_u = I up to the bases'
  orthogonality error), C = Qt D_i;
  m = n - r rows of C are free of couplings: F = C[r:], QR of F^T = H_P [Rq; 0];
  G = H_P^T C[:r]^T -> E = G[:m]^T (r x m), Dt = G[m:]^T (r x r);
  H_P^T V_i -> Vh = [:m] (m x s), Vt = [m:] (r x s).
Parent p of a, b: the reduced block, basis and column basis
  D_p = [[Dt_a, B_ab Vt_b^T], [B_ba Vt_a^T, Dt_b]],  U_p = R_p (transfer),
  V_p = [Vt_a W_p(a); Vt_b W_p(b)]       (n = r_a + r_b rows)
In the coordinates x_i = H_P [z_i; xt_i] the rows of node i read
  Rq^T z_i = (Qt beta_i)[r:],  Dt xt_i + f_i = bt_i = (Qt beta_i)[:r] - E z_i,
where f_i is the coupling coefficient of the HSS matvec (f_a = B_ab g_b +
R_p(a) f_p), so the second block is a reduced HSS system one level up.

Solve: beta_leaf = b(I_i); for a parent, beta_p = [bt_a - B_ab gk_b;
bt_b - B_ba gk_a] with gk_i = Vh^T z_i + W_i^T [gk_a; gk_b], the known part
of V_i^T x_i (the sibling's known unknowns enter the parent's rows through
B; unknowns outside the parent act through R_p, i.e. through U_p, and wait
for the parent's own bt).  Root: Rq^T z = beta_root.  Back substitution top
down: [xt_a; xt_b] = H_P [z_p; xt_p], x(I_i) = H_P [z_i; xt_i] at leaves.
Every transform is orthogonal; the only solves are triangular.
"""
import numpy as np
import scipy.linalg as sla


def _qr_complete(a):
    """Householder QR with LAPACK's sign convention: (Q n x n, R n x k)."""
    q, r = np.linalg.qr(a, mode="complete")
    return q, r


class HSSULV:
    """ULV factorisation of an oracle.hss_sampled.HSSSampled operator.

    skip_sibling=True leaves the sibling's known part (B_ab gk_b) out of the
    parent's right-hand side: a planted mistake for the tests only."""

    def __init__(self, hss, skip_sibling=False):
        self.h = h = hss
        self.skip_sibling = skip_sibling
        nodes = h.nodes
        nn = len(nodes)
        self.f = [None] * nn
        dmax = max(nd.depth for nd in nodes)
        for d in range(dmax, -1, -1):
            for i, nd in enumerate(nodes):
                if nd.depth != d:
                    continue
                root = i == 0
                if nd.left < 0:
                    D = h.D[i]
                    U = None if root else h.U[i]
                    V = None if root else h.V[i]
                else:
                    a, b = nd.left, nd.right
                    bab, bba = h.B[i]
                    fa, fb = self.f[a], self.f[b]
                    D = np.block([[fa["Dt"], bab @ fb["Vt"].T], [bba @ fa["Vt"].T, fb["Dt"]]])
                    U = None if root else h.R[i]
                    sa = h.V[a].shape[1]
                    V = None if root else np.vstack([fa["Vt"] @ h.W[i][:sa], fb["Vt"] @ h.W[i][sa:]])
                n = D.shape[0]
                if U is None:
                    r, qt = 0, np.eye(n)
                else:
                    r = U.shape[1]
                    if r > n:
                        raise ValueError("ULV needs r <= n at every node")
                    qu, ru = _qr_complete(U)
                    sg = np.where(np.diag(ru) < 0, -1.0, 1.0)
                    qt = qu.T.copy()
                    qt[:r] *= sg[:, None]
                c = qt @ D
                m = n - r
                hp, rq = _qr_complete(c[r:].T)
                g = hp.T @ c[:r].T
                fi = dict(n=n, r=r, m=m, Qt=qt, Hp=hp, Rq=rq[:m, :m], E=g[:m].T, Dt=g[m:].T)
                if V is not None:
                    hv = hp.T @ V
                    fi["Vh"], fi["Vt"] = hv[:m], hv[m:]
                self.f[i] = fi

    def solve(self, b):
        h, nodes = self.h, self.h.nodes
        b = np.asarray(b, dtype=float).reshape(h.N, -1)
        nr = b.shape[1]
        nn = len(nodes)
        z, bt, gk = [None] * nn, [None] * nn, [None] * nn
        dmax = max(nd.depth for nd in nodes)
        for d in range(dmax, -1, -1):
            for i, nd in enumerate(nodes):
                if nd.depth != d:
                    continue
                fi = self.f[i]
                if nd.left < 0:
                    beta = b[nd.start:nd.start + nd.len]
                else:
                    a, c = nd.left, nd.right
                    bab, bba = h.B[i]
                    k = 0.0 if self.skip_sibling else 1.0
                    beta = np.vstack([bt[a] - k * (bab @ gk[c]), bt[c] - k * (bba @ gk[a])])
                gam = fi["Qt"] @ beta
                r, m = fi["r"], fi["m"]
                z[i] = sla.solve_triangular(fi["Rq"], gam[r:], trans="T", lower=False) if m else np.zeros((0, nr))
                if i == 0:
                    continue
                bt[i] = gam[:r] - fi["E"] @ z[i]
                g = fi["Vh"].T @ z[i]
                if nd.left >= 0:
                    sa = h.V[nd.left].shape[1]
                    g = g + h.W[i][:sa].T @ gk[nd.left] + h.W[i][sa:].T @ gk[nd.right]
                gk[i] = g
        x = np.zeros((h.N, nr))
        xt = [None] * nn
        xt[0] = np.zeros((0, nr))
        for d in range(0, dmax + 1):
            for i, nd in enumerate(nodes):
                if nd.depth != d:
                    continue
                y = self.f[i]["Hp"] @ np.vstack([z[i], xt[i]])
                if nd.left < 0:
                    x[nd.start:nd.start + nd.len] = y
                else:
                    ra = self.f[nd.left]["r"]
                    xt[nd.left], xt[nd.right] = y[:ra], y[ra:]
        return x
