"""Seeded synthetic molecular system: orbital energies, Fock, antisymmetrized
integrals and dipole matrices, in the spin-orbital layout the reference uses.

The reference gets these from psi4 (CCSD_Helper.py:26-93 for TEI/F,
CCSD_Helper.py:970-980 for the dipoles).  psi4 is not available, so this module
is the integral provider.  Every value is a pure function of
(config id, component, canonical index) through a counter-based generator
(Philox4x32-10, Salmon et al., "Parallel random numbers: as easy as 1, 2, 3",
SC'11), so any block can be produced independently and identically on the host
(numpy, here) or on the device (the `tdcc_synth_tei` kernel in the HIP library).

Layout conventions (same as the reference):
  * spin orbitals, occupied first: p in [0, o) occupied, [o, o+v) virtual;
  * TEI[p,q,r,s] = <pq||rs>, real, with <pq||rs> = -<qp||rs> = <rs||pq>
    (hence also -<pq||sr>); entries with p == q or r == s are zero;
  * F real symmetric (n, n); mu real symmetric (3, n, n).

Distributions:
  * eps_occ ~ U[-1.5, -0.3] Eh and eps_vir ~ U[0.2, 2.0] Eh, each sorted ascending;
  * F = diag(eps) + symmetric off-diagonal N(0, (1e-3)^2);
  * one N(0, 0.05^2) draw per canonical orbit of the TEI symmetry group;
  * mu_x, mu_y, mu_z symmetric N(0, 0.3^2) (diagonal included).
"""
from __future__ import annotations

import os

import numpy as np

KEY0 = 0x20261203

# component ids of the counter
COMP_EPS_OCC = 0
COMP_EPS_VIR = 1
COMP_FOCK = 2
COMP_TEI = 3
COMP_MU = (4, 5, 6)
COMP_W0 = 7
COMP_AMP = 8  # 8..15: re/im of t1, t2, lam1, lam2 perturbations (test states)

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = 0x9E3779B9
_W1 = 0xBB67AE85
_MASK = np.uint64(0xFFFFFFFF)
_S32 = np.uint64(32)

TWO_PI = 6.283185307179586


def philox4x32_10(c0, c1, c2, c3, k0: int, k1: int):
    """Philox4x32 with 10 rounds on uint32 counter arrays (broadcastable).

    Returns four uint64 arrays holding the 32-bit outputs.
    """
    c0 = np.asarray(c0, dtype=np.uint64) & _MASK
    c1 = np.asarray(c1, dtype=np.uint64) & _MASK
    c2 = np.asarray(c2, dtype=np.uint64) & _MASK
    c3 = np.asarray(c3, dtype=np.uint64) & _MASK
    k0 = int(k0) & 0xFFFFFFFF
    k1 = int(k1) & 0xFFFFFFFF
    for r in range(10):
        if r:
            k0 = (k0 + _W0) & 0xFFFFFFFF
            k1 = (k1 + _W1) & 0xFFFFFFFF
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> _S32, p0 & _MASK
        hi1, lo1 = p1 >> _S32, p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ np.uint64(k0), lo1,
                          hi0 ^ c3 ^ np.uint64(k1), lo0)
    return c0, c1, c2, c3


def _u53(a, b):
    """Two 32-bit words -> double uniform in the open interval (0, 1)."""
    a = (a >> np.uint64(5)).astype(np.float64)
    b = (b >> np.uint64(6)).astype(np.float64)
    return (a * 67108864.0 + b + 0.5) * (1.0 / 9007199254740992.0)


def _counter(component: int, idx):
    idx = np.asarray(idx, dtype=np.uint64)
    return (np.full(idx.shape, component, dtype=np.uint64), idx & _MASK,
            idx >> _S32, np.zeros(idx.shape, dtype=np.uint64))


def uniform(config_id: int, component: int, idx):
    x0, x1, _, _ = philox4x32_10(*_counter(component, idx), KEY0, config_id)
    return _u53(x0, x1)


def normal(config_id: int, component: int, idx):
    """Box-Muller on the two uniforms of one Philox call (cosine branch)."""
    x0, x1, x2, x3 = philox4x32_10(*_counter(component, idx), KEY0, config_id)
    u1 = _u53(x0, x1)
    u2 = _u53(x2, x3)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(TWO_PI * u2)


def normals(config_id: int, component: int, size: int, chunk: int = 1 << 20, workers: int | None = None):
    """normal(config_id, component, 0..size-1), generated in chunks by a thread pool."""
    from concurrent.futures import ThreadPoolExecutor

    out = np.empty(size)

    def fill(start):
        stop = min(size, start + chunk)
        out[start:stop] = normal(config_id, component, np.arange(start, stop, dtype=np.uint64))

    if workers is None:
        workers = min(16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(max(1, workers)) as ex:
        list(ex.map(fill, range(0, size, chunk)))
    return out


def orbital_energies(config_id: int, o: int, v: int):
    eo = -1.5 + 1.2 * uniform(config_id, COMP_EPS_OCC, np.arange(o))
    ev = 0.2 + 1.8 * uniform(config_id, COMP_EPS_VIR, np.arange(v))
    return np.sort(eo), np.sort(ev)


def fock(config_id: int, o: int, v: int, offdiag_sigma: float = 1e-3):
    n = o + v
    eo, ev = orbital_energies(config_id, o, v)
    p, q = np.triu_indices(n, 1)
    off = offdiag_sigma * normal(config_id, COMP_FOCK, p * n + q)
    F = np.zeros((n, n))
    F[p, q] = off
    F[q, p] = off
    F[np.arange(n), np.arange(n)] = np.concatenate([eo, ev])
    return F


def tei_canonical(p, q, r, s, n: int):
    """Canonical orbit index and sign of <pq||rs> (sign 0 for p==q or r==s)."""
    p = np.asarray(p, dtype=np.int64)
    q = np.asarray(q, dtype=np.int64)
    r = np.asarray(r, dtype=np.int64)
    s = np.asarray(s, dtype=np.int64)
    sign = np.where((p == q) | (r == s), 0.0, 1.0)
    sw = p > q
    sign = np.where(sw, -sign, sign)
    p, q = np.where(sw, q, p), np.where(sw, p, q)
    sw = r > s
    sign = np.where(sw, -sign, sign)
    r, s = np.where(sw, s, r), np.where(sw, r, s)
    a = p * n + q
    b = r * n + s
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    return lo * (n * n) + hi, sign


def tei(config_id: int, o: int, v: int, sigma: float = 0.05, chunk: int = 1 << 20, workers: int | None = None):
    """Dense antisymmetrized TEI (n, n, n, n) float64.

    Chunks of the flat index are generated by a thread pool (numpy releases
    the GIL inside its array operations); every element is a pure function of
    its index, so the result does not depend on the chunking or the threads."""
    from concurrent.futures import ThreadPoolExecutor

    n = o + v
    out = np.empty(n ** 4, dtype=np.float64)
    total = n ** 4

    def fill(start):
        lin = np.arange(start, min(total, start + chunk), dtype=np.int64)
        s = lin % n
        r = (lin // n) % n
        q = (lin // (n * n)) % n
        p = lin // (n * n * n)
        idx, sign = tei_canonical(p, q, r, s, n)
        out[start:start + lin.size] = sign * (sigma * normal(config_id, COMP_TEI, idx))

    if workers is None:
        workers = min(16, len(os.sched_getaffinity(0)))
    starts = range(0, total, chunk)
    if workers <= 1 or total <= chunk:
        for st in starts:
            fill(st)
    else:
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(fill, starts))
    return out.reshape(n, n, n, n)


def dipoles(config_id: int, o: int, v: int, sigma: float = 0.3):
    n = o + v
    p, q = np.triu_indices(n, 0)
    mu = np.zeros((3, n, n))
    for c in range(3):
        d = sigma * normal(config_id, COMP_MU[c], p * n + q)
        mu[c, p, q] = d
        mu[c, q, p] = d
    return mu


def w0_draws(config_id: int, count: int, lo: float = 0.3, hi: float = 1.2):
    return lo + (hi - lo) * uniform(config_id, COMP_W0, np.arange(count))


def tei_block(config_id: int, o: int, v: int, classes: str, sigma: float = 0.05):
    """One block of the TEI, e.g. 'oovv', without building the full (n,n,n,n)."""
    n = o + v
    rng = {"o": np.arange(o), "v": np.arange(o, n)}
    axes = [rng[c] for c in classes]
    p, q, r, s = np.meshgrid(*axes, indexing="ij")
    idx, sign = tei_canonical(p, q, r, s, n)
    return sign * (sigma * normal(config_id, COMP_TEI, idx))


def mp2_t2(config_id: int, o: int, v: int, sigma: float = 0.05):
    """MP2 amplitudes <ij||ab> / (e_i + e_j - e_a - e_b) from the sorted
    orbital energies (the denominators of the reference's MP2_E,
    CCSD_Helper.py:102-124, with eigh(F) eigenvalues)."""
    eps = np.linalg.eigh(fock(config_id, o, v))[0]
    eo, ev = eps[:o], eps[o:]
    D = eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev
    return tei_block(config_id, o, v, "oovv", sigma) / D


def perturbed_state(config_id: int, o: int, v: int, sigma: float = 0.05, scale1: float = 0.02,
                    scale2: float = 0.005):
    """A seeded complex test state (t1, t2, lam1, lam2): t1, lam1 ~ scale1 (N + iN);
    t2, lam2 = MP2 amplitudes + scale2 (N + iN) antisymmetrized in (i, j) and
    (a, b).  Every element is a pure function of its index (components
    COMP_AMP..COMP_AMP+7), so host and tests on any machine build the same state."""
    mp2 = mp2_t2(config_id, o, v, sigma)
    out = []
    for k, shape in enumerate(((o, v), (o, o, v, v), (o, v), (o, o, v, v))):
        size = int(np.prod(shape))
        x = (normals(config_id, COMP_AMP + 2 * k, size) + 1j * normals(config_id, COMP_AMP + 2 * k + 1, size))
        x = x.reshape(shape)
        if len(shape) == 2:
            out.append(scale1 * x)
        else:
            x = x - x.transpose(1, 0, 2, 3)
            x = x - x.transpose(0, 1, 3, 2)
            out.append(mp2 + (0.25 * scale2) * x)
    return tuple(out)


class SyntheticSystem:
    """System provider for the propagation path: F, mu and the TEI of one
    (config id, o, v).

    With `tei_on_device=True` the (n,n,n,n) TEI is not built on the host;
    the HIP library generates it from `tei_config = (config_id, sigma)`.
    """

    def __init__(self, config_id: int, o: int, v: int, with_tei: bool = True,
                 tei_sigma: float = 0.05, tei_on_device: bool = False):
        self.config_id = config_id
        self.o = self.nocc = o
        self.v = self.nvir = v
        self.n = o + v
        self.tei_sigma = tei_sigma
        self.F = fock(config_id, o, v)
        self.mu = dipoles(config_id, o, v)
        self.tei_config = (config_id, tei_sigma)
        self.TEI = tei(config_id, o, v, sigma=tei_sigma) if (with_tei and not tei_on_device) else None

    def tei_block(self, classes: str):
        if self.TEI is not None:
            sl = {"o": slice(0, self.o), "v": slice(self.o, self.n)}
            return self.TEI[tuple(sl[c] for c in classes)]
        return tei_block(self.config_id, self.o, self.v, classes, self.tei_sigma)
