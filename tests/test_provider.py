"""The integral provider (python_code_to_check_c_amd/provider.py) on the CPU.

Pins: the spin-orbital construction of the reference's psi4 path
(CCSD_Helper.py:26-93, :970-980) is checked against textbook closed-shell
formulas that do not go through spin orbitals at all: the Fock matrix
F_PQ = h_PQ + sum_M^docc [2 (PQ|MM) - (PM|MQ)] on both spins, and the
Hartree-Fock energy sum_i h_ii + 1/2 sum_ij <ij||ij> = 2 sum_I h_II +
sum_IJ [2 (II|JJ) - (IJ|JI)]; the integrals vanish between different spins and
are antisymmetric; FCIDUMP files round-trip; malformed systems are rejected.
"""
import numpy as np
import pytest

from python_code_to_check_c_amd import provider


def spatial_system(N=5, ndocc=2, seed=4):
    rng = np.random.default_rng(seed)
    h = 0.1 * rng.standard_normal((N, N))
    h = h + h.T + np.diag(np.sort(rng.uniform(-2.0, 1.0, N)))
    h[np.arange(ndocc), np.arange(ndocc)] -= 8.0  # occupied well below the virtual orbitals
    B = 0.3 * rng.standard_normal((6, N, N))
    B = B + B.transpose(0, 2, 1)
    eri = np.einsum("Lpq,Lrs->pqrs", B, B)
    mu = 0.3 * rng.standard_normal((3, N, N))
    mu = mu + mu.transpose(0, 2, 1)
    return h, eri, mu, ndocc


def test_spin_orbital_construction_matches_closed_shell_formulas():
    h, eri, mu, nd = spatial_system()
    N = h.shape[0]
    s = provider.from_spatial(h, eri, mu, nd)
    assert (s.nocc, s.nvir) == (2 * nd, 2 * N - 2 * nd)
    f_spatial = h + 2.0 * np.einsum("pqmm->pq", eri[:, :, :nd, :nd]) - np.einsum("pmmq->pq", eri[:, :nd, :nd, :])
    F = s.F
    for a in range(2):
        assert np.abs(F[a::2, a::2] - f_spatial).max() <= 1e-13
    assert np.abs(F[0::2, 1::2]).max() == 0.0
    o = s.nocc
    e_spin = np.trace(provider.spin_tile(h)[:o, :o]) + 0.5 * np.einsum("ijij->", s.TEI[:o, :o, :o, :o])
    e_closed = (2.0 * np.trace(h[:nd, :nd]) + 2.0 * np.einsum("iijj->", eri[:nd, :nd, :nd, :nd])
                - np.einsum("ijji->", eri[:nd, :nd, :nd, :nd]))
    assert abs(e_spin - e_closed) <= 1e-12
    for k in range(3):
        assert np.array_equal(s.mu[k][0::2, 0::2], mu[k]) and np.array_equal(s.mu[k][1::2, 1::2], mu[k])


def test_spin_integrals_structure():
    _, eri, _, _ = spatial_system()
    T = provider.spin_tei(eri)
    n = T.shape[0]
    spin = np.arange(n) % 2
    p, q, r, s = np.meshgrid(*(np.arange(n),) * 4, indexing="ij")
    allowed = ((spin[p] == spin[r]) & (spin[q] == spin[s])) | ((spin[p] == spin[s]) & (spin[q] == spin[r]))
    assert np.abs(T[~allowed]).max() == 0.0
    assert np.abs(T + T.transpose(1, 0, 2, 3)).max() <= 1e-15
    assert np.abs(T - T.transpose(2, 3, 0, 1)).max() <= 1e-15
    # same-spin element: <pq||rs> = (PR|QS) - (PS|QR) for p, q, r, s all alpha
    assert T[0, 2, 4, 6] == eri[0, 2, 1, 3] - eri[0, 3, 1, 2]


def test_fcidump_and_npz_round_trip(tmp_path):
    h, eri, mu, nd = spatial_system()
    path = str(tmp_path / "FCIDUMP")
    provider.write_fcidump(path, h, eri, 2 * nd, ecore=1.25)
    h2, eri2, nelec, ecore = provider.read_fcidump(path)
    assert nelec == 2 * nd and ecore == 1.25
    assert np.array_equal(h2, h) and np.abs(eri2 - eri).max() <= 1e-15
    a = provider.load(path, mu=mu)
    b = provider.from_spatial(h, eri, mu, nd)
    np.savez(tmp_path / "spatial.npz", h=h, eri=eri, mu=mu, ndocc=nd)
    c = provider.load(str(tmp_path / "spatial.npz"))
    np.savez(tmp_path / "spin.npz", TEI=b.TEI, F=b.F, mu=b.mu, nocc=b.nocc)
    d = provider.load(str(tmp_path / "spin.npz"))
    for x in (a, c, d):
        assert x.nocc == b.nocc and np.abs(x.TEI - b.TEI).max() <= 1e-15 and np.abs(x.F - b.F).max() <= 1e-15
        assert np.array_equal(x.mu, b.mu)


def test_malformed_systems_are_rejected(tmp_path):
    h, eri, mu, nd = spatial_system()
    b = provider.from_spatial(h, eri, mu, nd)
    bad = b.TEI.copy()
    bad[0, 1, 2, 3] += 1e-3
    with pytest.raises(ValueError, match="antisymmetr"):
        provider.ArraySystem(b.nocc, b.F, b.mu, bad)
    F = b.F.copy()
    F[0, 3] += 1e-3
    with pytest.raises(ValueError, match="symmetric"):
        provider.ArraySystem(b.nocc, F, b.mu, b.TEI)
    F = b.F.copy()
    F[0, 0] = F.diagonal().max() + 1.0  # an occupied orbital above every virtual one
    with pytest.raises(ValueError, match="occupied orbitals must come first"):
        provider.ArraySystem(b.nocc, F, b.mu, b.TEI)
    e = eri.copy()
    e[0, 1, 2, 3] += 1e-3
    with pytest.raises(ValueError, match="8-fold"):
        provider.from_spatial(h, e, mu, nd)
    path = str(tmp_path / "FCIDUMP")
    provider.write_fcidump(path, h, eri, 2 * nd)
    with pytest.raises(ValueError, match="dipole"):
        provider.load(path)
