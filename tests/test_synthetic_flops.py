"""The input generator and the FLOP model (no GPU)."""
import numpy as np
import pytest

from oracle.flops import step_flops
from python_code_to_check_c_amd import synthetic


def test_philox_known_answers():
    """Philox4x32-10 known-answer vectors of Salmon et al. (Random123 kat_vectors)."""
    cases = [
        ((0, 0, 0, 0, 0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 6, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for args, want in cases:
        got = synthetic.philox4x32_10(*args)
        assert tuple(int(x) for x in got) == want


def test_tei_symmetries_and_blocks():
    o, v = 3, 5
    T = synthetic.tei(7, o, v, sigma=0.05)
    assert np.array_equal(T, -T.transpose(1, 0, 2, 3))
    assert np.array_equal(T, -T.transpose(0, 1, 3, 2))
    assert np.array_equal(T, T.transpose(2, 3, 0, 1))
    n = o + v
    for p in range(n):
        assert not np.any(T[p, p]) and not np.any(T[:, :, p, p])
    for key in ("oovv", "vovv", "vvvv", "ovov"):
        sl = {"o": slice(0, o), "v": slice(o, n)}
        assert np.array_equal(synthetic.tei_block(7, o, v, key), T[tuple(sl[c] for c in key)])
    assert 0.04 < synthetic.tei(8, 6, 10).std() < 0.06


def test_fock_and_dipoles():
    F = synthetic.fock(2, 10, 38)
    eo, ev = np.diag(F)[:10], np.diag(F)[10:]
    assert np.all(np.diff(eo) >= 0) and np.all(np.diff(ev) >= 0)
    assert eo.min() >= -1.5 and eo.max() <= -0.3 and ev.min() >= 0.2 and ev.max() <= 2.0
    assert np.array_equal(F, F.T)
    mu = synthetic.dipoles(2, 10, 38)
    assert np.array_equal(mu, mu.transpose(0, 2, 1))


@pytest.mark.parametrize("method,ov,f_ref", [
    ("CCSD", (10, 4), 6.131e7), ("CCSD", (10, 38), 5.774e10), ("CCSD", (20, 160), 5.895e13),
    ("CCSD", (20, 96), 8.40e12), ("CC2", (10, 38), 1.706e10), ("CC2", (20, 160), 1.555e13)])
def test_flops_as_written(method, ov, f_ref):
    """F_ref per RK4 step agrees with the independent static count of the reference source."""
    r, a = step_flops(*ov, method=method)
    assert abs(r / f_ref - 1) < 2e-3
    assert a < r


def test_flops_hoisted_c3():
    r, a = step_flops(20, 160, "CCSD")
    assert abs(a / 3.166e13 - 1) < 5e-3
