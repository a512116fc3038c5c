"""System providers for real molecules: user-supplied integrals, Fock and dipoles.

The reference builds its system from psi4 (CCSD_Helper.py:26-93, :970-980):
  * TEI = mints.mo_spin_eri(C, C): antisymmetrized physicist's integrals
    <pq||rs> = <pq|rs> - <pq|sr> over spin orbitals, spin orbital p = 2 P + s
    (spatial orbital P, s = 0 alpha / 1 beta, interleaved), with
    <pq|rs> = (PR|QS) when spin(p) = spin(r) and spin(q) = spin(s), else 0;
  * F = H + sum over occupied m of <pm||qm>, H the spatial core Hamiltonian in
    the MO basis tiled over the two spins (F_MO, :82-93);
  * dipoles: the spatial MO dipole matrices tiled over the two spins
    (Defd_dipole, :970-980);
  * occupied spin orbitals first: occ = slice(2 ndocc) (:49-50).
This module builds the same arrays from spatial MO integrals given directly, in
an .npz file or in an FCIDUMP file (with the dipoles beside it), or takes
spin-orbital arrays as they are; either way the result is checked for the
layout the kernels assume (antisymmetry, real symmetric F and dipoles,
occupied orbitals below the virtual ones) before it is handed out.

A provider has the attributes the drop-in classes read: nocc, nvir, F (n,n),
mu (3,n,n) and TEI (n,n,n,n).
"""
from __future__ import annotations

import re

import numpy as np


class ArraySystem:
    """Spin-orbital system: nocc occupied spin orbitals first, TEI = <pq||rs>."""

    def __init__(self, nocc, F, mu, TEI, tol=1e-10, check=True):
        self.F = np.ascontiguousarray(F, dtype=np.float64)
        self.mu = np.ascontiguousarray(mu, dtype=np.float64)
        self.TEI = np.ascontiguousarray(TEI, dtype=np.float64)
        n = self.F.shape[0]
        self.nocc = self.o = int(nocc)
        self.nvir = self.v = n - self.nocc
        self.n = n
        if check:
            check_system(self.nocc, self.F, self.mu, self.TEI, tol)


def check_system(nocc, F, mu, TEI, tol=1e-10):
    """Raise ValueError unless the arrays have the layout the propagation path
    assumes; the tolerance is relative to each array's largest element."""
    F, mu, TEI = np.asarray(F), np.asarray(mu), np.asarray(TEI)
    n = F.shape[0] if F.ndim == 2 else -1
    if F.shape != (n, n) or mu.shape != (3, n, n) or TEI.shape != (n, n, n, n):
        raise ValueError(f"shapes F {F.shape}, mu {mu.shape}, TEI {TEI.shape} do not describe one system")
    if not 0 < nocc < n:
        raise ValueError(f"nocc = {nocc} must lie in (0, {n})")
    for name, x in (("F", F), ("mu", mu), ("TEI", TEI)):
        if not np.all(np.isfinite(x)):
            raise ValueError(f"{name} has non-finite entries")
    scale = lambda x: max(float(np.abs(x).max()), 1e-300)
    if np.abs(F - F.T).max() > tol * scale(F):
        raise ValueError("F is not symmetric")
    for k in range(3):
        if np.abs(mu[k] - mu[k].T).max() > tol * scale(mu[k]):
            raise ValueError(f"dipole matrix {'xyz'[k]} is not symmetric")
    s = tol * scale(TEI)
    if np.abs(TEI + TEI.transpose(1, 0, 2, 3)).max() > s or np.abs(TEI + TEI.transpose(0, 1, 3, 2)).max() > s:
        raise ValueError("TEI is not antisymmetrized: <pq||rs> = -<qp||rs> = -<pq||sr> must hold")
    if np.abs(TEI - TEI.transpose(2, 3, 0, 1)).max() > s:
        raise ValueError("TEI lacks the symmetry <pq||rs> = <rs||pq> of real orbitals")
    d = np.diag(F)
    if d[:nocc].max() >= d[nocc:].min():
        raise ValueError("the occupied orbitals must come first: max diag F over occupied "
                         f"{d[:nocc].max():.6g} >= min over virtual {d[nocc:].min():.6g}")


def spin_tile(x):
    """Spatial (N,N) matrix -> spin orbitals (2N,2N), interleaved alpha/beta, as
    the reference tiles H and the dipoles (repeat + identity tile)."""
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[0]
    return x.repeat(2, axis=1).repeat(2, axis=0) * np.tile(np.identity(2), (N, N))


def spin_tei(eri):
    """Spatial chemist's integrals (PQ|RS) (N,N,N,N) -> antisymmetrized
    physicist's spin-orbital integrals <pq||rs> (2N,2N,2N,2N), interleaved spins."""
    eri = np.asarray(eri, dtype=np.float64)
    N = eri.shape[0]
    phys = eri.transpose(0, 2, 1, 3)  # <PQ|RS> = (PR|QS)
    s = np.arange(2 * N) % 2
    same = (s[:, None] == s[None, :]).astype(np.float64)  # delta(spin p, spin r)
    big = np.kron(phys, np.ones((2, 2, 2, 2)))
    direct = big * same[:, None, :, None] * same[None, :, None, :]
    return direct - direct.transpose(0, 1, 3, 2)


def from_spatial(h, eri, mu, ndocc, tol=1e-10):
    """Provider from spatial MO integrals: core Hamiltonian h (N,N), chemist's
    electron repulsion integrals (PQ|RS) (N,N,N,N), dipoles mu (3,N,N), ndocc
    doubly occupied orbitals (the first ndocc in the given order)."""
    h = np.asarray(h, dtype=np.float64)
    eri = np.asarray(eri, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    N = h.shape[0]
    if h.shape != (N, N) or eri.shape != (N,) * 4 or mu.shape != (3, N, N):
        raise ValueError(f"shapes h {h.shape}, eri {eri.shape}, mu {mu.shape} do not describe one system")
    sc = max(float(np.abs(eri).max()), 1e-300)
    for perm in ((1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1)):
        if np.abs(eri - eri.transpose(perm)).max() > tol * sc:
            raise ValueError("eri lacks the 8-fold symmetry of real chemist's integrals (PQ|RS)")
    TEI = spin_tei(eri)
    occ = slice(0, 2 * int(ndocc))
    F = spin_tile(h) + np.einsum("pmqm->pq", TEI[:, occ, :, occ])
    muS = np.stack([spin_tile(mu[k]) for k in range(3)])
    return ArraySystem(2 * int(ndocc), F, muS, TEI, tol=tol)


def read_fcidump(path):
    """FCIDUMP text: header &FCI NORB=..., NELEC=..., ... &END, then lines
    'value i j k l' (1-based, chemist's (ij|kl), each symmetry-unique element
    once), 'value i j 0 0' for h_ij and 'value 0 0 0 0' for the core energy.
    Returns (h, eri, nelec, ecore)."""
    with open(path) as f:
        text = f.read()
    m = re.search(r"&FCI(.*?)(&END|/)", text, flags=re.S | re.I)
    if not m:
        raise ValueError(f"{path}: no &FCI ... &END header")
    head = m.group(1)

    def field(name):
        mm = re.search(name + r"\s*=\s*(-?\d+)", head, flags=re.I)
        if not mm:
            raise ValueError(f"{path}: header lacks {name}")
        return int(mm.group(1))

    N, nelec = field("NORB"), field("NELEC")
    h = np.zeros((N, N))
    eri = np.zeros((N, N, N, N))
    ecore = 0.0
    for ln, line in enumerate(text[m.end():].splitlines(), 1):
        parts = line.split()
        if not parts:
            continue
        if len(parts) != 5:
            raise ValueError(f"{path}: integral line {ln} has {len(parts)} fields")
        val = float(parts[0].replace("D", "E").replace("d", "e"))
        i, j, k, l = (int(x) for x in parts[1:])
        if not all(0 <= x <= N for x in (i, j, k, l)):
            raise ValueError(f"{path}: integral line {ln} has an index outside 0..{N}")
        if i == j == k == l == 0:
            ecore = val
        elif k == 0 and l == 0:
            h[i - 1, j - 1] = h[j - 1, i - 1] = val
        else:
            p, q, r, s = i - 1, j - 1, k - 1, l - 1
            for a, b, c, d in ((p, q, r, s), (q, p, r, s), (p, q, s, r), (q, p, s, r),
                               (r, s, p, q), (s, r, p, q), (r, s, q, p), (s, r, q, p)):
                eri[a, b, c, d] = val
    return h, eri, nelec, ecore


def write_fcidump(path, h, eri, nelec, ecore=0.0):
    """FCIDUMP text of spatial integrals (each 8-fold symmetry-unique (ij|kl) once)."""
    N = h.shape[0]
    lines = [f" &FCI NORB={N},NELEC={nelec},MS2=0,", "  ORBSYM=" + "1," * N, "  ISYM=1,", " &END"]
    for i in range(N):
        for j in range(i + 1):
            for k in range(N):
                for l in range(k + 1):
                    if i * (i + 1) // 2 + j < k * (k + 1) // 2 + l:
                        continue
                    lines.append(f"{float(eri[i, j, k, l])!r} {i + 1} {j + 1} {k + 1} {l + 1}")
    for i in range(N):
        for j in range(i + 1):
            lines.append(f"{float(h[i, j])!r} {i + 1} {j + 1} 0 0")
    lines.append(f"{float(ecore)!r} 0 0 0 0")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def load(path, mu=None, tol=1e-10):
    """Provider from a file.

    * .npz with spin-orbital arrays: TEI (n,n,n,n) = <pq||rs>, F (n,n), mu (3,n,n), nocc;
    * .npz with spatial MO arrays: h (N,N), eri (N,N,N,N) chemist's, mu (3,N,N), ndocc;
    * FCIDUMP text (any other file): h and eri from it, ndocc = NELEC / 2 (closed
      shell), dipoles `mu` (3,N,N) passed in (FCIDUMP carries none).
    """
    if str(path).endswith(".npz"):
        with np.load(path, allow_pickle=False) as z:
            keys = set(z.files)
            if {"TEI", "F", "mu", "nocc"} <= keys:
                return ArraySystem(int(z["nocc"]), z["F"], z["mu"], z["TEI"], tol=tol)
            if {"h", "eri", "mu", "ndocc"} <= keys:
                return from_spatial(z["h"], z["eri"], z["mu"], int(z["ndocc"]), tol=tol)
            raise ValueError(f"{path}: expected keys TEI, F, mu, nocc or h, eri, mu, ndocc; found {sorted(keys)}")
    h, eri, nelec, _ = read_fcidump(path)
    if nelec % 2:
        raise ValueError(f"{path}: NELEC = {nelec} is odd; only closed shells are supported")
    if mu is None:
        raise ValueError("an FCIDUMP file carries no dipole integrals: pass mu (3, NORB, NORB)")
    return from_spatial(h, eri, mu, nelec // 2, tol=tol)
