"""Oracle pins: two-electron CCSD == FCI, the RK driver on a linear mock, and
the restart file layout."""
import cmath
import os

import numpy as np
import pandas as pd
import pytest

from oracle import fci, ground_state, rk4
from oracle.cc_equations import CCEquations
from python_code_to_check_c_amd import restart_io, synthetic


# ------------------------------------------------------------------ o = 2 FCI
@pytest.mark.parametrize("cid,v", [(31, 5), (32, 4)])
def test_two_electron_ccsd_is_fci(cid, v):
    s = synthetic.SyntheticSystem(cid, 2, v)
    eq = CCEquations(s.TEI, 2)
    t1, t2, _ = ground_state.solve_t(eq, s.F, tol=1e-12)
    l1, l2, _ = ground_state.solve_lambda(eq, s.F, t1, t2, tol=1e-12)
    E, Eref, gamma = fci.two_electron_fci(s.F, s.TEI)
    assert abs(eq.CCSD_Corr_E(t1, t2, s.F) - (E - Eref)) < 1e-11
    D = eq.density(t1, t2, l1, l2)
    occ = np.zeros_like(gamma)
    occ[0, 0] = occ[1, 1] = 1.0
    assert np.abs(D + occ - gamma).max() < 1e-10
    # dipole of the correlated density equals the FCI dipole minus the reference one
    mu = eq.dipole_moment(t1, t2, l1, l2, s.mu)
    for c in range(3):
        assert abs(mu[c] - np.sum(s.mu[c] * (gamma - occ))) < 1e-10


# ------------------------------------------------------------ RK linear mock
class LinearMock:
    """R(y) = M y + c for every chain; records the F each stage sees."""

    def __init__(self, rng, o=2, v=3):
        self.o, self.v = o, v
        self.M1 = 0.3 * (rng.standard_normal((o * v, o * v)) + 1j * rng.standard_normal((o * v, o * v)))
        self.M2 = 0.3 * (rng.standard_normal((o * o * v * v,) * 2) + 1j * rng.standard_normal((o * o * v * v,) * 2))
        self.c1 = rng.standard_normal(o * v) + 1j * rng.standard_normal(o * v)
        self.c2 = rng.standard_normal(o * o * v * v) + 0j
        self.seen = []

    def _r1(self, y, F):
        self.seen.append(F)
        return (self.M1 @ y.ravel() + self.c1).reshape(y.shape)

    def _r2(self, y, F):
        return (self.M2 @ y.ravel() + self.c2).reshape(y.shape)

    def T1eq_rhs(self, t1, t2, F):
        return self._r1(t1, F)

    def T2eq_rhs(self, t1, t2, F):
        return self._r2(t2, F)

    def lam_1eq_rhs(self, t1, t2, l1, l2, F):
        return self._r1(l1, F)

    def lam2eq_rhs(self, t1, t2, l1, l2, F):
        return self._r2(l2, F)


def _phi(M, h, r):
    out, term = np.zeros_like(r), r.copy()
    fact = 1.0
    for k in range(4):
        fact *= (k + 1)
        out += h * term / fact
        term = h * (M @ term)
    return out


def test_rk_step_closed_form():
    rng = np.random.default_rng(3)
    m = LinearMock(rng)
    y1 = rng.standard_normal((2, 3)) + 1j * rng.standard_normal((2, 3))
    y2 = rng.standard_normal((2, 2, 3, 3)) + 1j * rng.standard_normal((2, 2, 3, 3))
    dt, t = 0.01, 0.37
    Vt = lambda s: np.zeros((5, 5))
    new, base = rk4.rk4_step(m, "CCSD", np.zeros((5, 5)), y1, y2, y1, y2, t, dt, Vt, 15)
    r1 = m.M1 @ y1.ravel() + m.c1
    r2 = m.M2 @ y2.ravel() + m.c2
    exp_t1 = np.around(y1, 15) - 1j * _phi(m.M1, dt, r1).reshape(2, 3)
    exp_l1 = np.around(y1, 15) + 1j * _phi(m.M1, dt, r1).reshape(2, 3)
    exp_t2 = np.around(y2, 15) - 1j * _phi(m.M2, dt, r2).reshape(y2.shape)
    assert np.abs(new[0] - exp_t1).max() < 1e-13
    assert np.abs(new[1] - exp_t2).max() < 1e-12
    assert np.abs(new[2] - exp_l1).max() < 1e-13
    for b, y in zip(base, (y1, y2, y1, y2)):
        assert np.array_equal(b.real, np.rint(y.real * 1e15) / 1e15)
        assert np.array_equal(b.imag, np.rint(y.imag * 1e15) / 1e15)


def test_rk_stage_times_and_field():
    """Stages see V(t), V(t+dt/2) twice, V(t+dt) with t already advanced; V is
    -A mu exp(+i 2 pi w0 t) (CCSD_Helper.py:1197-1210)."""
    rng = np.random.default_rng(4)
    m = LinearMock(rng)
    mu = rng.standard_normal((3, 5, 5))
    y1 = np.zeros((2, 3), complex)
    y2 = np.zeros((2, 2, 3, 3), complex)
    A, w0, dt, t0 = 0.002, 0.923, 0.001, 0.0
    m.dipole_moment = lambda a, b, c, d, mu_: [0j, 0j, 0j]
    rk4.propagate(m, np.zeros((5, 5)), y1, y2, y1, y2, mu, w0, A, t0, dt, max_steps=2)
    t = t0 + dt
    expect = [t, t + dt / 2, t + dt / 2, t + dt]
    for k, s in enumerate(expect):
        V = -A * mu[2] * cmath.exp(1j * w0 * 2 * np.pi * s)
        assert np.array_equal(m.seen[k], np.zeros((5, 5)) + V)
    assert len(m.seen) == 2 * 8  # T1 and L1 chains, 4 stages each, 2 steps


def test_rk_divergence_stop_saves_rounded_base(tmp_path):
    rng = np.random.default_rng(6)
    m = LinearMock(rng)
    calls = {"n": 0}

    def dip(a, b, c, d, mu_):
        calls["n"] += 1
        return [0j, 0j, (1000.0 if calls["n"] >= 3 else 0.0) + 0j]

    m.dipole_moment = dip
    y1 = rng.standard_normal((2, 3)) + 0j
    y2 = np.zeros((2, 2, 3, 3), complex)
    seen = []
    tr = rk4.propagate(m, np.zeros((5, 5)), y1, y2, y1, y2, np.zeros((3, 5, 5)), 0.9, 0.0, 0.0, 0.01,
                       max_steps=10, observe=lambda i, t, s: seen.append(s))
    assert tr.stop_reason == "divergence" and tr.steps == 2
    assert np.array_equal(tr.saved[0], np.around(seen[0][0], 15))
    assert abs(tr.saved_time - 0.01) < 1e-15


# ------------------------------------------------------------- restart files
def test_restart_layout_roundtrip(tmp_path):
    rng = np.random.default_rng(8)
    o, v = 2, 3
    t1 = rng.standard_normal((o, v)) + 1j * rng.standard_normal((o, v))
    t2 = rng.standard_normal((o, o, v, v)) + 1j * rng.standard_normal((o, o, v, v))
    t1[0, 0] = -1e-17 + 1.5e-5j
    F = rng.standard_normal((o + v, o + v))
    data = pd.DataFrame([[0.001, 0.5, -0.25]], index=[1], columns=("time", "mu_real", "mu_imag"))
    timing = pd.DataFrame([[1.0, 0.1, 0.2, 0.3, 0.4]], index=[1], columns=("total", "t1", "t2", "l1", "l2"))
    restart_io.save_data(F, t1, t2, t1, t2, data, timing, 15, None, path=str(tmp_path))
    restart_io.save_parameters(0.923, 0.002, 0.0123, 0.001, 15, o, v, "CCSD", path=str(tmp_path))
    lines = open(tmp_path / "t1_real.dat").read().splitlines()
    assert lines[0] == "0\t0\t-0.0"
    assert lines[1] == f"0\t1\t{np.around(t1[0, 1].real, 15)}"
    assert open(tmp_path / "t1_imag.dat").read().splitlines()[0] == "0\t0\t1.5e-05"
    assert open(tmp_path / "t2_real.dat").read().splitlines()[1].startswith("0\t0\t0\t1\t")
    assert open(tmp_path / "Parameters.csv").read().splitlines() == [
        ",w0,A,t0,dt,precs,i,a,CCSD_or_CC2",
        "1,0.9230000000,0.0020000000,0.0123000000,0.0010000000,15,2,3,CCSD"]
    assert open(tmp_path / "H2O.csv").read().splitlines()[0] == ",time,mu_real,mu_imag"
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        r = restart_io.read_restart()
    finally:
        os.chdir(cwd)
    assert r["method"] == "CCSD" and (r["o"], r["v"]) == (o, v)
    assert abs(r["t0"] - (0.0123 + 0.001)) < 1e-15 and abs(r["tf"] - 50.0123) < 1e-12
    assert np.array_equal(r["t1"], np.around(t1, 15))
    assert np.array_equal(r["t2"], np.around(t2, 15))
    assert np.array_equal(r["F"], np.around(F, 15))
