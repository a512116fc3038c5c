"""The ctypes wrappers reject wrongly shaped host arrays before any C call
(the C side reads fixed sizes from every host pointer).  CPU only: the
context object is made without creating a device context."""
import numpy as np
import pytest

from python_code_to_check_c_amd import _lib


def _ctx(o=3, v=4):
    c = object.__new__(_lib.Context)
    c.o, c.v, c.n = o, v, o + v
    c._h = None
    return c


def _amps(o, v):
    return np.zeros((o, v), complex), np.zeros((o, o, v, v), complex)


def test_rhs_rejects_bad_shapes():
    c = _ctx()
    t1, t2 = _amps(3, 4)
    F = np.zeros((7, 7), complex)
    with pytest.raises(ValueError):
        c.rhs(_lib.RHS_T1, np.zeros((6, 6)), t1, t2)
    with pytest.raises(ValueError):
        c.rhs(_lib.RHS_T2, F, t1[:, :3], t2)
    with pytest.raises(ValueError):
        c.rhs(_lib.RHS_L1, F, t1, t2, t1, t2[:2])
    with pytest.raises(ValueError):
        c.rhs(_lib.RHS_L2, F, t1, t2)  # Lambda residuals need lam1 and lam2
    with pytest.raises(ValueError):
        c.rhs(_lib.RHS_T1, F, t1, t2, t1, None)


def test_intermediate_and_dipole_reject_bad_shapes():
    c = _ctx()
    t1, t2 = _amps(3, 4)
    F = np.zeros((7, 7), complex)
    with pytest.raises(ValueError):
        c.intermediate("Fae", (4, 4), np.zeros((7, 6)), t1, t2, t1, t2)
    with pytest.raises(ValueError):
        c.intermediate("Fae", (4, 4), F, t1, t2, t1[:2], t2)
    with pytest.raises(ValueError):
        c.dipole(t1, t2, t1, t2.reshape(3, 3, 16))
    with pytest.raises(ValueError):
        c.set_state(t1, t2, t1.T, t2)
