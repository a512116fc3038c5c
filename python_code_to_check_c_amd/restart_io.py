"""Restart and output files in the reference's layout.

Writers follow CCSD_Helper.py:1102-1145, readers CC_Calculator.py:70-135:
  Parameters.csv        pandas frame ``,w0,A,t0,dt,precs,i,a,CCSD_or_CC2`` with
                        row index 1, floats ``%.10f``; t0 holds the stop time t - dt
  {F,t1,lam1}_{real,imag}.dat   ``i<TAB>a<TAB>value`` per element, C order
  {t2,lam2}_{real,imag}.dat     ``i<TAB>j<TAB>a<TAB>b<TAB>value``
                        value = str(numpy.around(x, precs)), csv module, '\\n' endings
  H2O.csv               ``,time,mu_real,mu_imag`` (appended without header on restart)
  timing.csv            ``,total,t1,t2,l1,l2``
  restart_state.bin     exact binary snapshot (state after the last step, its
                        time, F, field parameters) for a bit-for-bit resume;
                        written beside the text files, read only on request
"""
from __future__ import annotations

import os

import numpy as np
import pandas as pd

from . import _lib


def write_2data(F, FileName, precs):
    """CCSD_Helper.py:1107-1114 (rows i, a, numpy.around(value, precs)),
    written by the C++ writer of the library."""
    _lib.write_restart_block(np.asarray(F, dtype=np.float64), FileName, precs)


def write_4data(F, FileName, precs):
    """CCSD_Helper.py:1116-1125 (rows i, j, a, b, numpy.around(value, precs)),
    written by the C++ writer of the library."""
    _lib.write_restart_block(np.asarray(F, dtype=np.float64), FileName, precs)


def save_parameters(w0, A, t, dt, precs, i, a, method, path="."):
    """CCSD_Helper.py:1102-1105 (the t0 column receives the time passed in)."""
    df = pd.DataFrame(columns=("w0", "A", "t0", "dt", "precs", "i", "a", "CCSD_or_CC2"))
    df.loc[1] = [w0, A, t, dt, precs, i, a, method]
    df.to_csv(os.path.join(path, "Parameters.csv"), float_format="%.10f")


def save_data(F, t1, t2, lam1, lam2, data, timing, precs, restart, path="."):
    """CCSD_Helper.py:1127-1145"""
    j = lambda name: os.path.join(path, name)
    if restart is None:
        data.to_csv(j("H2O.csv"))
        timing.to_csv(j("timing.csv"))
    else:
        with open(j("H2O.csv"), "a") as f:
            data.to_csv(f, header=False)
        with open(j("timing.csv"), "a") as f:
            timing.to_csv(f, header=False)
    F = np.asarray(F)
    write_2data(F.real, j("F_real.dat"), precs)
    write_2data(F.imag if np.iscomplexobj(F) else np.zeros_like(F), j("F_imag.dat"), precs)
    write_2data(t1.real, j("t1_real.dat"), precs)
    write_2data(t1.imag, j("t1_imag.dat"), precs)
    write_4data(t2.real, j("t2_real.dat"), precs)
    write_4data(t2.imag, j("t2_imag.dat"), precs)
    write_2data(lam1.real, j("lam1_real.dat"), precs)
    write_2data(lam1.imag, j("lam1_imag.dat"), precs)
    write_4data(lam2.real, j("lam2_real.dat"), precs)
    write_4data(lam2.imag, j("lam2_imag.dat"), precs)


STATE_FILE = "restart_state.bin"


def convert_2data(FileName, i, a):
    """CC_Calculator.py:85-96: rows (i, a, value) -> dense (i, a), by the
    library's C++ reader."""
    return _lib.read_restart_block(FileName, (i, a))


def convert_4data(FileName, i, a):
    """CC_Calculator.py:98-113: rows (i, j, a, b, value) -> dense (i, i, a, a),
    by the library's C++ reader."""
    return _lib.read_restart_block(FileName, (i, i, a, a))


def save_state(F, t1, t2, lam1, lam2, t, dt, w0, A, precs, method, path="."):
    """Exact snapshot beside the text files: the state after the stopping step
    and its time t (not the rounded base and t - dt the text files hold)."""
    _lib.write_state_file(os.path.join(path, STATE_FILE), method, precs, t, dt, w0, A, F, t1, t2, lam1, lam2)


def read_state(path="."):
    """The exact snapshot as a dict (method, precs, t, dt, w0, A, F, t1, t2, lam1, lam2)."""
    return _lib.read_state_file(os.path.join(path, STATE_FILE))


def read_parameters(path="."):
    """Parameters.csv as TDCC_restart reads it (CC_Calculator.py:71-83):
    w0, A, t0 (advanced by dt), tf = 50 + the saved t0, dt, precs, method, o, v."""
    p = pd.read_csv(os.path.join(path, "Parameters.csv"))
    t0 = p["t0"][0]
    dt = p["dt"][0]
    return dict(w0=p["w0"][0], A=p["A"][0], t0=t0 + dt, tf=50.0 + t0, dt=dt, precs=int(p["precs"][0]),
                method=p["CCSD_or_CC2"][0], o=int(p["i"][0]), v=int(p["a"][0]))


def read_restart(path="."):
    """Everything TDCC_restart reads (CC_Calculator.py:70-135).

    Returns a dict with w0, A, t0 (already advanced by dt, :80), tf (= 50 + the
    saved t0, :77), dt, precs, method, o, v, F (real part only, :135), t1, t2,
    lam1, lam2.
    """
    j = lambda name: os.path.join(path, name)
    prm = read_parameters(path)
    w0, A, t0, tf, dt, precs, method = (prm[k] for k in ("w0", "A", "t0", "tf", "dt", "precs", "method"))
    i, a = prm["o"], prm["v"]
    t1 = convert_2data(j("t1_real.dat"), i, a) + 1.0 * 1j * convert_2data(j("t1_imag.dat"), i, a)
    lam1 = convert_2data(j("lam1_real.dat"), i, a) + 1.0 * 1j * convert_2data(j("lam1_imag.dat"), i, a)
    t2 = convert_4data(j("t2_real.dat"), i, a) + 1.0 * 1j * convert_4data(j("t2_imag.dat"), i, a)
    lam2 = convert_4data(j("lam2_real.dat"), i, a) + 1.0 * 1j * convert_4data(j("lam2_imag.dat"), i, a)
    F = convert_2data(j("F_real.dat"), i + a, i + a)
    return dict(w0=w0, A=A, t0=t0, tf=tf, dt=dt, precs=precs, method=method, o=i, v=a, F=F,
                t1=t1, t2=t2, lam1=lam1, lam2=lam2)
