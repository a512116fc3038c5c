"""The library's restart block writer against the reference's csv writer.

oracle/restart_text.py restates write_2data / write_4data of
CCSD_Helper.py:1107-1125.  It is pinned here by known answers of Python's
float printing and numpy's rounding (including the exponent switches at
1e-4 / 1e16, negative zero and values that round to zero), and by reading
every written value back with float().  The C++ writer must produce the same
bytes for random and edge-case blocks at several precs.
"""
import numpy as np
import pytest

from oracle import restart_text
from python_code_to_check_c_amd import _lib

EDGE = np.array([0.0, -0.0, 0.1, -0.1, 1.0, 2.5, 1e-5, 1.5e-5, -2.5e-7, 1e-4, 9.99e-5, 0.000123456789,
                 123.456, 1e15, 1e16, -1.5e16, 123456789012345.6, 1.2345678901234568e17, 3e-16, -4e-16,
                 0.3333333333333333, -0.6666666666666666, 1e-300, 5e-324, 7.0e22])


def _bytes(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("value,precs,text", [
    (0.1, 15, "0.1"), (1e-5, 15, "1e-05"), (1e-4, 15, "0.0001"), (-0.0, 15, "-0.0"),
    (-1e-20, 15, "-0.0"), (1e16, 15, "1e+16"), (1e15, 15, "1000000000000000.0"),
    # numpy.around scales, rounds half to even, scales back: 2.675 * 100 is
    # 267.5 in binary, so 2.68 (Python's round() would give 2.67)
    (2.675, 2, "2.68"), (0.125, 2, "0.12"), (1.5, 0, "2.0"), (2.5, 0, "2.0"),
    (1.23456789e-7, 10, "1.235e-07"), (-1.5e-5, 15, "-1.5e-05"),
])
def test_oracle_known_answers(tmp_path, value, precs, text):
    p = tmp_path / "k.dat"
    restart_text.write_2data(np.array([[value]]), str(p), precs)
    assert _bytes(p) == f"0\t0\t{text}\n".encode()


def test_oracle_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3, 4, 5, 6)) * 10.0 ** rng.integers(-9, 9, (3, 4, 5, 6))
    p = tmp_path / "r.dat"
    restart_text.write_4data(x, str(p), 15)
    rows = [line.split("\t") for line in p.read_text().splitlines()]
    assert len(rows) == x.size
    back = np.array([float(r[4]) for r in rows]).reshape(x.shape)
    assert np.array_equal(back, np.around(x, 15))
    idx = np.array([[int(v) for v in r[:4]] for r in rows])
    assert np.array_equal(idx, np.array(list(np.ndindex(x.shape))))


@pytest.mark.parametrize("precs", [15, 10, 3, 0])
def test_writer_matches_reference_2d(tmp_path, precs):
    rng = np.random.default_rng(precs)
    x = np.concatenate([EDGE, rng.standard_normal(40) * 10.0 ** rng.integers(-12, 18, 40)]).reshape(5, 13)
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    restart_text.write_2data(x, str(a), precs)
    _lib.write_restart_block(x, str(b), precs)
    assert _bytes(a) == _bytes(b)


@pytest.mark.parametrize("precs", [15, 12])
def test_writer_matches_reference_4d(tmp_path, precs):
    rng = np.random.default_rng(10 + precs)
    x = rng.standard_normal((3, 3, 7, 7)) * 0.05
    x = x - x.transpose(1, 0, 2, 3)
    x.flat[:EDGE.size] = EDGE
    a, b = tmp_path / "a.dat", tmp_path / "b.dat"
    restart_text.write_4data(x, str(a), precs)
    _lib.write_restart_block(x, str(b), precs)
    assert _bytes(a) == _bytes(b)


def test_writer_errors(tmp_path):
    with pytest.raises(_lib.TdccError):
        _lib.write_restart_block(np.zeros((2, 2)), str(tmp_path / "no" / "such" / "dir.dat"), 15)
    with pytest.raises(ValueError):
        _lib.write_restart_block(np.zeros(3), str(tmp_path / "x.dat"), 15)
