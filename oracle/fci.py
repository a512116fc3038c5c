"""Brute-force FCI for two electrons in n spin orbitals (test infrastructure).

Used to pin the CCSD equations at their fixed point: for two electrons CCSD is
exact, so the converged CCSD correlation energy equals E_FCI - E_ref, and the
Lambda-based correlated 1-RDM of CCSD_Helper.py:984-1006 plus the reference
occupation equals the FCI 1-RDM.

Inputs are the spin-orbital Fock matrix F and antisymmetrized <pq||rs>; the
core Hamiltonian is h = F - sum_m <pm||qm> over the two occupied orbitals
(the inverse of F_MO, CCSD_Helper.py:92).
"""
from __future__ import annotations

import itertools

import numpy as np


def two_electron_fci(F, TEI, nocc=2):
    assert nocc == 2
    n = F.shape[0]
    h = F - np.einsum("pmqm->pq", TEI[:, :2, :, :2])
    dets = list(itertools.combinations(range(n), 2))
    index = {d: k for k, d in enumerate(dets)}
    H = np.zeros((len(dets), len(dets)))
    for a, (p, q) in enumerate(dets):
        for b, (r, s) in enumerate(dets):
            v = TEI[p, q, r, s]
            v += h[p, r] * (q == s) - h[p, s] * (q == r) - h[q, r] * (p == s) + h[q, s] * (p == r)
            H[a, b] = v
    e_ref = H[index[(0, 1)], index[(0, 1)]]
    w, X = np.linalg.eigh(H)
    # the state the CC solution is connected to: largest weight on the reference
    k = int(np.argmax(np.abs(X[index[(0, 1)], :])))
    c = X[:, k]
    gamma = np.zeros((n, n))
    for b, (r, s) in enumerate(dets):
        for v_, (rest, sign_v) in ((r, (s, 1.0)), (s, (r, -1.0))):
            for u in range(n):
                if u == rest:
                    continue
                det, sgn = ((u, rest), 1.0) if u < rest else ((rest, u), -1.0)
                gamma[u, v_] += c[index[det]] * c[b] * sign_v * sgn
    return w[k], e_ref, gamma
