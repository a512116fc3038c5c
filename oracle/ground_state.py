"""Ground-state T and Lambda amplitudes for the propagation inputs (test infrastructure).

The reference converges them with a Jacobi update preconditioned by eigh(F)
denominators (CCSD_Helper.py:324-344, :658-677) accelerated by DIIS
(:347-416, :702-776).  Only the fixed point matters to the propagation path,
so this solver uses the same Jacobi step with a standard Pulay DIIS on the
amplitude differences and stops on the residual max-abs instead of the energy
difference.
"""
from __future__ import annotations

import numpy as np


class NotConverged(RuntimeError):
    pass


def denominators(F, o):
    eps = np.linalg.eigh(F)[0]
    eo, ev = eps[:o], eps[o:]
    d1 = eo.reshape(-1, 1) - ev
    d2 = eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev
    return d1, d2


class _Diis:
    def __init__(self, size):
        self.size = size
        self.vecs = []
        self.errs = []

    def push(self, vec, err):
        self.vecs.append(vec.copy())
        self.errs.append(err.copy())
        if len(self.vecs) > self.size:
            self.vecs.pop(0)
            self.errs.pop(0)

    def extrapolate(self):
        m = len(self.vecs)
        if m < 2:
            return self.vecs[-1]
        B = -np.ones((m + 1, m + 1), dtype=complex)
        B[-1, -1] = 0
        for a in range(m):
            for b in range(m):
                B[a, b] = np.vdot(self.errs[a], self.errs[b])
        rhs = np.zeros(m + 1, dtype=complex)
        rhs[-1] = -1
        c = np.linalg.lstsq(B, rhs, rcond=None)[0]
        return sum(c[k] * self.vecs[k] for k in range(m))


def _solve(residual, unpack, pack, x, d, tol, maxiter, diis_size, label):
    diis = _Diis(diis_size)
    rmax = np.inf
    for it in range(maxiter):
        r = residual(x)
        rmax = np.abs(r).max()
        if rmax < tol:
            return x, it, rmax
        xn = x + r / d
        diis.push(xn, xn - x)
        x = diis.extrapolate()
        if not np.all(np.isfinite(x)):
            break
    raise NotConverged(f"{label}: residual max-abs {rmax:.3e} after {maxiter} iterations")


def solve_t(eq, F, method="CCSD", tol=1e-10, maxiter=200, diis_size=8):
    """Return (t1, t2, iterations) with max|R1|, max|R2| < tol."""
    o, v = eq.nocc, eq.nvir
    d1, d2 = denominators(F, o)
    n1 = o * v
    t2rhs = eq.T2eq_rhs if method == "CCSD" else eq.T2eq_rhs_CC2

    def pack(t1, t2):
        return np.concatenate([t1.ravel(), t2.ravel()]).astype(complex)

    def unpack(x):
        return x[:n1].reshape(o, v), x[n1:].reshape(o, o, v, v)

    def residual(x):
        t1, t2 = unpack(x)
        return pack(eq.T1eq_rhs(t1, t2, F), t2rhs(t1, t2, F))

    oovv = eq.blk("oovv")
    x0 = pack(np.zeros((o, v)), oovv / d2)
    x, it, _ = _solve(residual, unpack, pack, x0, pack(d1, d2), tol, maxiter,
                      diis_size, f"{method} T")
    t1, t2 = unpack(x)
    return t1, t2, it


def solve_lambda(eq, F, t1, t2, method="CCSD", tol=1e-10, maxiter=200, diis_size=8):
    """Return (lam1, lam2, iterations) with max|L1|, max|L2| < tol."""
    o, v = eq.nocc, eq.nvir
    d1, d2 = denominators(F, o)
    n1 = o * v
    if method == "CCSD":
        l1rhs, l2rhs = eq.lam_1eq_rhs, eq.lam2eq_rhs
    else:
        l1rhs, l2rhs = eq.L1eq_rhs_CC2, eq.L2eq_rhs_CC2

    def pack(a, b):
        return np.concatenate([a.ravel(), b.ravel()]).astype(complex)

    def unpack(x):
        return x[:n1].reshape(o, v), x[n1:].reshape(o, o, v, v)

    def residual(x):
        l1, l2 = unpack(x)
        return pack(l1rhs(t1, t2, l1, l2, F), l2rhs(t1, t2, l1, l2, F))

    x, it, _ = _solve(residual, unpack, pack, pack(t1, t2), pack(d1, d2), tol,
                      maxiter, diis_size, f"{method} Lambda")
    l1, l2 = unpack(x)
    return l1, l2, it
