"""The drop-in Python surface on the GPU, checked against the oracle.

CCSD_Helper.Runge_Kutta_solver / CC2_Helper.Runge_Kutta_solver_CC2 with the
reference's signature (CCSD_Helper.py:1191, CC2_Helper.py:437) and the
stop / restart cycle of CC_Calculator (CC_Calculator.py:56-145): a run that
reaches tf writes nothing; a run stopped by the wall-clock rule writes the
restart files with the rounded base of its last step, and TDCC_restart
resumes from them.  Inputs: the golden o=4, v=6 system and amplitudes.
Tolerance 1e-11 on dipoles; restart amplitudes agree to the 15-decimal
rounding (2e-15).
"""
import os
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import rk4
from oracle.cc_equations import CCEquations
from python_code_to_check_c_amd import restart_io
from python_code_to_check_c_amd.CC2_Helper import CC2_Helper
from python_code_to_check_c_amd.CC_Calculator import CC_Calculator
from python_code_to_check_c_amd.CCSD_Helper import CCSD_Helper

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
W0, A, DT = 0.923, 0.002, 0.001


@pytest.fixture(scope="module")
def g():
    return dict(np.load(os.path.join(HERE, "golden", "golden_o4v6.npz")))


def provider(g):
    return SimpleNamespace(nocc=int(g["o"]), nvir=int(g["v"]), F=g["F"], mu=g["mu"], TEI=g["TEI"])


def amplitudes(g):
    return g["t1"], g["t2"], g["lam1"], g["lam2"]


@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_solver_runs_to_tf_and_saves_nothing(g, method, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    print(f"\n[dropin to-tf {method}]", flush=True)
    mol = (CCSD_Helper if method == "CCSD" else CC2_Helper)(provider(g))
    mol.verbose = False
    solver = mol.Runge_Kutta_solver if method == "CCSD" else mol.Runge_Kutta_solver_CC2
    solver(g["F"], *amplitudes(g), W0, A, 0.0, 0.0125, DT, 100.0, 15)
    tr = rk4.propagate(CCEquations(g["TEI"], int(g["o"])), g["F"], *amplitudes(g), g["mu"], W0, A, 0.0, DT,
                       method=method, tf=0.0125)
    assert len(mol.data) == tr.steps == 13
    assert list(mol.data["time"]) == tr.times
    assert np.abs(mol.mu_history - np.array(tr.mu)).max() <= 1e-11
    assert mol.stop_reason is None
    assert os.listdir(tmp_path) == []


def test_timeout_stop_then_restart(g, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    print("\n[dropin stop/restart CCSD]", flush=True)
    o = int(g["o"])
    eq = CCEquations(g["TEI"], o)
    calc = CC_Calculator(provider(g), w0=W0, A=A, t0=0.0, dt=DT, precs=15)
    mol = calc.mol
    mol.verbose = False
    # timeout 0: the wall-clock rule fires at the end of the first chunk (one step)
    mol.Runge_Kutta_solver(g["F"], *amplitudes(g), W0, A, 0.0, 10.1, DT, 0.0, 15)
    assert mol.stop_reason == "timeout"
    n1 = len(mol.data)
    tr = rk4.propagate(eq, g["F"], *amplitudes(g), g["mu"], W0, A, 0.0, DT, method="CCSD", max_steps=n1)
    assert np.abs(mol.mu_history - np.array(tr.mu)).max() <= 1e-11
    for f in ("Parameters.csv", "t1_real.dat", "t1_imag.dat", "t2_real.dat", "t2_imag.dat",
              "lam1_real.dat", "lam1_imag.dat", "lam2_real.dat", "lam2_imag.dat", "F_real.dat", "H2O.csv"):
        assert os.path.exists(f), f
    # the saved state is the rounded base of the last step taken
    before = rk4.propagate(eq, g["F"], *amplitudes(g), g["mu"], W0, A, 0.0, DT, method="CCSD",
                           max_steps=n1 - 1).state if n1 > 1 else amplitudes(g)
    r = restart_io.read_restart()
    for got, ref in zip((r["t1"], r["t2"], r["lam1"], r["lam2"]), before):
        assert np.abs(got - np.around(ref, 15)).max() <= 2e-15
    # resume: the restarted run starts from the files and stops again after one chunk
    calc.TDCC_restart(0.0)
    mol2 = calc.last
    n2 = len(mol2.data)
    tr2 = rk4.propagate(eq, r["F"], r["t1"], r["t2"], r["lam1"], r["lam2"], g["mu"], r["w0"], r["A"], r["t0"],
                        r["dt"], method="CCSD", max_steps=n2)
    assert np.abs(mol2.mu_history - np.array(tr2.mu)).max() <= 1e-11


def test_divergence_stop_saves_rounded_base(g, tmp_path, monkeypatch):
    """A = 100 drives |Re mu_z| above 100 at step 22, inside the fifth chunk
    (steps 16..31): the run is replayed up to that step, and the files hold the
    rounded base of step 22 and the time t - dt, as the reference saves them
    (CCSD_Helper.py:1248-1251)."""
    monkeypatch.chdir(tmp_path)
    print("\n[dropin divergence stop CCSD]", flush=True)
    o = int(g["o"])
    eq = CCEquations(g["TEI"], o)
    big_A = 100.0
    mol = CCSD_Helper(provider(g))
    mol.verbose = False
    mol.Runge_Kutta_solver(g["F"], *amplitudes(g), W0, big_A, 0.0, 10.1, DT, 1000.0, 15)
    tr = rk4.propagate(eq, g["F"], *amplitudes(g), g["mu"], W0, big_A, 0.0, DT, method="CCSD", tf=10.1)
    assert tr.stop_reason == "divergence" and tr.steps == 22
    assert mol.stop_reason == "divergence"
    assert len(mol.data) == 22
    scale = np.abs(np.array(tr.mu)).max()
    assert np.abs(mol.mu_history - np.array(tr.mu)).max() <= 1e-10 * scale
    r = restart_io.read_restart()
    assert abs(r["t0"] - DT - tr.saved_time) <= 1e-10  # the reader adds dt back (CC_Calculator.py:80)
    for got, ref in zip((r["t1"], r["t2"], r["lam1"], r["lam2"]), tr.saved):
        assert np.abs(got - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())


def test_exact_resume_is_bitwise_continuous(g, tmp_path, monkeypatch):
    """A run stopped by the wall-clock rule also writes the exact snapshot
    (restart_state.bin: the state after the stopping step and its time);
    resuming from it continues the interrupted run bit for bit, through
    Runge_Kutta_solver and through CC_Calculator.TDCC_restart(exact=True)."""
    monkeypatch.chdir(tmp_path)
    print("\n[dropin exact resume CCSD]", flush=True)
    tf = 0.0205
    full = CCSD_Helper(provider(g))
    full.verbose = False
    full.Runge_Kutta_solver(g["F"], *amplitudes(g), W0, A, 0.0, tf, DT, 1000.0, 15)
    full_mu, full_state = full.mu_history.copy(), full.get_state()
    assert full.stop_reason is None and len(full_mu) == 21
    calc = CC_Calculator(provider(g), w0=W0, A=A, t0=0.0, dt=DT, precs=15)
    first = calc.mol
    first.verbose = False
    first.Runge_Kutta_solver(g["F"], *amplitudes(g), W0, A, 0.0, tf, DT, 0.0, 15)
    assert first.stop_reason == "timeout"
    n1 = len(first.data)
    r = restart_io.read_state()
    assert r["t"] == first.final_time and r["method"] == "CCSD"
    rest = CCSD_Helper(provider(g))
    rest.verbose = False
    rest.Runge_Kutta_solver(r["F"], r["t1"], r["t2"], r["lam1"], r["lam2"], r["w0"], r["A"], r["t"], tf, r["dt"],
                            1000.0, r["precs"], "restart")
    mu = np.concatenate([first.mu_history, rest.mu_history])
    assert len(mu) == len(full_mu) and np.array_equal(mu, full_mu)
    for x, y in zip(rest.get_state(), full_state):
        assert np.array_equal(x, y)
    # the calculator's exact resume (timeout 0: one more step) gives step n1 + 1 of the full run
    calc.TDCC_restart(0.0, exact=True)
    assert calc.last.stop_reason == "timeout"
    assert np.array_equal(calc.last.mu_history[0], full_mu[n1])
