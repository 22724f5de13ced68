"""NumPy restatement of the AMG V-cycle the library applies (test
infrastructure): the hierarchy comes from the library's host set-up
(cfsa_ns.HostAmg); every smoothing step, restriction, coarse solve and
prolongation is written out here with scipy sparse products.

V-cycle on level l for right-hand side b (x0 = 0):
  pre-smoothing, degree-2 Chebyshev in D^-1 A:
    x1 = c0 D^-1 b
    x2 = x1 + c1 x1 + c2 D^-1 (b - A x1)
  b_c = R (b - A x2); e = V(b_c) on the next level (dense inverse on the last)
  x3 = x2 + P e
  post-smoothing from x3:
    x4 = x3 + c0 D^-1 (b - A x3)
    x5 = x4 + c1 (x4 - x3) + c2 D^-1 (b - A x4)
"""
import numpy as np
import scipy.sparse as sp


def csr(rp, ci, v, shape):
    return sp.csr_matrix((v, ci, rp), shape=shape)


class VCycle:
    def __init__(self, H, A0=None):
        L, nc, cjac = H.levels()
        self.levels = []
        for l in range(L):
            rp, ci, v, shape = H.matrix(l, "A")
            A = csr(rp, ci, v, shape) if (l > 0 or A0 is None) else A0
            prm = H.params(l)
            d = A.diagonal()
            dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
            lv = {"A": A, "dinv": dinv, "c": prm["coef"]}
            if l < L - 1:
                lv["P"] = csr(*H.matrix(l, "P")[:3], H.matrix(l, "P")[3])
                lv["R"] = csr(*H.matrix(l, "R")[:3], H.matrix(l, "R")[3])
            self.levels.append(lv)
        self.cjac = cjac
        self.cinv = None if cjac else H.coarseInverse()

    def apply(self, b, l=0):
        lv = self.levels[l]
        A, dinv = lv["A"], lv["dinv"]
        if l == len(self.levels) - 1:
            return dinv * b if self.cjac else self.cinv @ b
        c0, c1, c2 = lv["c"]
        x1 = c0 * (dinv * b)
        x2 = (x1 + c1 * x1) + c2 * (dinv * (b - A @ x1))
        e = self.apply(lv["R"] @ (b - A @ x2), l + 1)
        x3 = x2 + lv["P"] @ e
        x4 = x3 + c0 * (dinv * (b - A @ x3))
        return (x4 + c1 * (x4 - x3)) + c2 * (dinv * (b - A @ x4))


def pcg(A, b, prec, tol, maxit=5000):
    """CG with preconditioner prec, stopping on the Jacobi-scaled residual
    |D^-1 r| <= tol |D^-1 b| (the penalty rows dominate |b| itself)."""
    d = A.diagonal()
    x = np.zeros_like(b)
    r = b.copy()
    nb = np.linalg.norm(b / d)
    z = prec(r)
    p = z.copy()
    rho = r @ z
    for it in range(1, maxit + 1):
        q = A @ p
        al = rho / (p @ q)
        x += al * p
        r -= al * q
        if np.linalg.norm(r / d) <= tol * nb:
            return x, it
        z = prec(r)
        rn = r @ z
        p = z + (rn / rho) * p
        rho = rn
    return x, maxit
