"""Ground-state T and Lambda amplitudes on the GPU (inputs of the propagation).

The reference converges them with a Jacobi update preconditioned by eigh(F)
denominators (CCSD_Helper.py:324-344, :658-677) plus DIIS (:347-416,
:702-776; CC2_Helper.py:179-248, :324-400).  Here every residual is evaluated
by the HIP library through the helper's *_rhs methods; the Jacobi step and a
Pulay DIIS on the amplitude differences run on the host, and convergence is
judged on the residual max-abs.  The MP2 starting guess follows MP2_E
(CCSD_Helper.py:102-124).
"""
from __future__ import annotations

import numpy as np


class NotConverged(RuntimeError):
    pass


def denominators(F, o):
    eps = np.linalg.eigh(np.asarray(F).real)[0]
    eo, ev = eps[:o], eps[o:]
    d1 = eo.reshape(-1, 1) - ev
    d2 = eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev
    return d1, d2


def mp2_guess(helper, F):
    _, d2 = denominators(F, helper.nocc)
    return helper._oovv() / d2


def _diis_solve(residual, x0, d, tol, maxiter, size, label, verbose):
    vecs, errs = [], []
    x = x0
    rmax = np.inf
    for it in range(maxiter):
        r = residual(x)
        rmax = float(np.abs(r).max())
        if verbose:
            print(f"{label} iter {it:3d}  max|R| = {rmax:.3e}")
        if rmax < tol:
            return x, it
        xn = x + r / d
        vecs.append(xn)
        errs.append(xn - x)
        if len(vecs) > size:
            vecs.pop(0)
            errs.pop(0)
        m = len(vecs)
        if m < 2:
            x = xn
        else:
            B = -np.ones((m + 1, m + 1), complex)
            B[-1, -1] = 0
            for a in range(m):
                for b in range(a, m):
                    B[a, b] = np.vdot(errs[a], errs[b])
                    B[b, a] = np.conj(B[a, b])
            rhs = np.zeros(m + 1, complex)
            rhs[-1] = -1
            c = np.linalg.lstsq(B, rhs, rcond=None)[0]
            x = sum(c[k] * vecs[k] for k in range(m))
        if not np.all(np.isfinite(x)):
            break
    raise NotConverged(f"{label}: residual max-abs {rmax:.3e} after {maxiter} iterations")


def solve_t(helper, F, tol=1e-10, maxiter=200, diis_size=8, verbose=False):
    o, v = helper.nocc, helper.nvir
    d1, d2 = denominators(F, o)
    n1 = o * v
    t2rhs = helper.T2eq_rhs_CC2 if helper.method == "CC2" else helper.T2eq_rhs
    pack = lambda a, b: np.concatenate([a.ravel(), b.ravel()]).astype(complex)
    unpack = lambda x: (x[:n1].reshape(o, v), x[n1:].reshape(o, o, v, v))
    Fc = np.asarray(F, dtype=complex)

    def residual(x):
        t1, t2 = unpack(x)
        return pack(helper.T1eq_rhs(t1, t2, Fc), t2rhs(t1, t2, Fc))

    x0 = pack(np.zeros((o, v)), mp2_guess(helper, F))
    x, it = _diis_solve(residual, x0, pack(d1, d2), tol, maxiter, diis_size, helper.method + " T", verbose)
    t1, t2 = unpack(x)
    return t1, t2, it


def solve_lambda(helper, F, t1, t2, tol=1e-10, maxiter=200, diis_size=8, verbose=False):
    o, v = helper.nocc, helper.nvir
    d1, d2 = denominators(F, o)
    n1 = o * v
    if helper.method == "CC2":
        l1rhs, l2rhs = helper.L1eq_rhs_CC2, helper.L2eq_rhs_CC2
    else:
        l1rhs, l2rhs = helper.lam_1eq_rhs, helper.lam2eq_rhs
    pack = lambda a, b: np.concatenate([a.ravel(), b.ravel()]).astype(complex)
    unpack = lambda x: (x[:n1].reshape(o, v), x[n1:].reshape(o, o, v, v))
    Fc = np.asarray(F, dtype=complex)

    def residual(x):
        l1, l2 = unpack(x)
        return pack(l1rhs(t1, t2, l1, l2, Fc), l2rhs(t1, t2, l1, l2, Fc))

    x, it = _diis_solve(residual, pack(t1, t2), pack(d1, d2), tol, maxiter, diis_size,
                        helper.method + " Lambda", verbose)
    l1, l2 = unpack(x)
    return l1, l2, it
