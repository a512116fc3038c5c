"""The library's restart block reader and exact state file (no device).

The reader replaces convert_2data / convert_4data (CC_Calculator.py:85-113),
which read the blocks with pandas.read_csv and assign row by row into a zero
array.  Checked here: a block written by the library's writer (byte-identical
to the reference's csv writer, test_restart_writer.py) reads back to exactly
the rounded values (correctly rounded parsing); pandas.read_csv gives the same
values with its round-trip parser, and values at most 1 ulp away with its
default parser (the reference's), which is not correctly rounded for some
inputs (e.g. -304337750958.48895 reads as -304337750958.4889); rows in any order, missing rows (zeros), duplicate rows (the last one
wins), blank lines, CRLF endings and indices written as floats behave as in the
reference's loop; malformed rows and out-of-range indices are errors.  The
state file round-trips bit for bit and detects corruption.
"""
import numpy as np
import pandas as pd
import pytest

from python_code_to_check_c_amd import _lib

EDGE = np.array([0.0, -0.0, 0.1, -0.1, 1.0, 2.5, 1e-5, 1.5e-5, -2.5e-7, 1e-4, 9.99e-5, 0.000123456789,
                 123.456, 1e15, 1e16, -1.5e16, 123456789012345.6, 1.2345678901234568e17, 3e-16, -4e-16,
                 0.3333333333333333, -0.6666666666666666, 1e-300, 5e-324, 7.0e22])

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(_lib.LIB_PATH), reason="libtdcc_hip.so not built")


def _around(x, precs):
    return np.around(x, precs)


@pytest.mark.parametrize("shape", [(5, 7), (3, 3, 4, 4)])
@pytest.mark.parametrize("precs", [15, 8])
def test_round_trip_and_pandas(tmp_path, shape, precs):
    rng = np.random.default_rng(5)
    x = rng.standard_normal(shape) * 10.0 ** rng.integers(-18, 18, size=shape)
    x.reshape(-1)[:min(x.size, EDGE.size)] = EDGE[:x.size]
    p = str(tmp_path / "b.dat")
    _lib.write_restart_block(x, p, precs)
    got = _lib.read_restart_block(p, shape)
    ref = _around(x, precs)
    assert np.array_equal(got, ref) and np.array_equal(np.signbit(got), np.signbit(ref))
    names = ["i", "a", "F"] if len(shape) == 2 else ["i", "j", "a", "b", "F"]
    for prec in ("round_trip", None):
        df = pd.read_csv(p, sep="\t", header=None, names=names, float_precision=prec)
        pd_vals = np.zeros(shape)
        pd_vals[tuple(np.around(df[k].to_numpy()).astype(int) for k in names[:-1])] = df["F"].to_numpy()
        if prec == "round_trip":
            assert np.array_equal(got, pd_vals)
        else:  # pandas' default parser is not always correctly rounded: at most 1 ulp off
            assert np.all(np.abs(got - pd_vals) <= np.spacing(np.abs(got)))


def test_reference_loop_semantics(tmp_path):
    p = tmp_path / "r.dat"
    # out of order, a missing element, a duplicate (last wins), float indices,
    # a blank line and CRLF endings
    p.write_bytes(b"1\t2\t0.5\r\n0\t0\t-1.25\n\n1.0\t2.0\t7.0\n0\t1\t1e-05\n")
    got = _lib.read_restart_block(str(p), (2, 3))
    ref = np.zeros((2, 3))
    ref[0, 0], ref[0, 1], ref[1, 2] = -1.25, 1e-05, 7.0
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("text", [b"0\t0\n", b"0\t5\t1.0\n", b"0\t0\t1.0\t2.0\n", b"0\tx\t1.0\n", b"-1\t0\t1.0\n"])
def test_malformed_rows_are_errors(tmp_path, text):
    p = tmp_path / "bad.dat"
    p.write_bytes(text)
    with pytest.raises(_lib.TdccError):
        _lib.read_restart_block(str(p), (2, 3))


def test_state_file_round_trip_and_checksum(tmp_path):
    o, v = 3, 4
    rng = np.random.default_rng(9)
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    F = rng.standard_normal((o + v, o + v))
    st = (c(o, v), c(o, o, v, v), c(o, v), c(o, o, v, v))
    p = str(tmp_path / "restart_state.bin")
    _lib.write_state_file(p, "CC2", 15, 0.1234567890123456789, 0.001, 0.923, 0.002, F, *st)
    r = _lib.read_state_file(p)
    assert r["method"] == "CC2" and r["precs"] == 15
    assert (r["t"], r["dt"], r["w0"], r["A"]) == (0.1234567890123456789, 0.001, 0.923, 0.002)
    assert np.array_equal(r["F"], F)
    for k, x in zip(("t1", "t2", "lam1", "lam2"), st):
        assert np.array_equal(r[k], x), k
    raw = bytearray(open(p, "rb").read())
    raw[200] ^= 1
    open(p, "wb").write(bytes(raw))
    with pytest.raises(_lib.TdccError):
        _lib.read_state_file(p)
