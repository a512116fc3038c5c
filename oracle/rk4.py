"""The reference's partial-RK4 propagator, restated in numpy (test infrastructure).

Follows CCSD_Helper.py:1158-1256 (CCSD) and CC2_Helper.py:412-500 (CC2):
  * four decoupled chains: T1 varies only t1, T2 only t2, L1 only lam1, L2
    only lam2; the other amplitudes stay at the step-start *unrounded* values
    (CCSD_Helper.py:1216-1225);
  * each chain's stages feed the raw right-hand side (no -i/+i inside);
    the chain returns dt/6 (k1 + 2k2 + 2k3 + k4) (:1158-1187);
  * the increment is multiplied by -1j (T) or +1j (Lambda) and added to the
    base rounded with numpy.around(x, precs) (:1207-1214, :1230-1233);
  * time is advanced before the right-hand sides are evaluated (:1210);
  * the field is V(t) = -A mu_pol exp(i 2 pi w0 t), complex (:1197-1200);
  * the dipole is taken after the update (:1238); CCSD labels time with
    round(t, 4), CC2 with round(t, 5) (:1239, CC2_Helper.py:484);
  * stop rules: wall clock > 0.9 timeout*60 s, or |Re mu_z| > 100; on a stop
    the saved state is the rounded base of the stopping step and the saved
    time is t - dt (:1242-1251).  Running to t >= tf saves nothing.
"""
from __future__ import annotations

import cmath
import time

import numpy as np


def field(mu_pol, A, w0, t):
    """V(t) exactly as CCSD_Helper.py:1197-1200 evaluates it."""
    return -A * mu_pol * cmath.exp(1j * w0 * 2 * np.pi * t)


def chain(rhs, y, t, dt, F, Vt):
    """One partial-RK4 chain over y (CCSD_Helper.py:1158-1187)."""
    k1 = rhs(y, F + Vt(t))
    k2 = rhs(y + dt / 2.0 * k1, F + Vt(t + dt / 2.0))
    k3 = rhs(y + dt / 2.0 * k2, F + Vt(t + dt / 2.0))
    k4 = rhs(y + dt * k3, F + Vt(t + dt))
    return dt / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


def rhs_set(eq, method):
    """The four right-hand sides used by the method's RK loop.

    CC2 keeps the CCSD T1 equation (CC2_Helper.py:463).
    """
    if method == "CCSD":
        return (eq.T1eq_rhs, eq.T2eq_rhs, eq.lam_1eq_rhs, eq.lam2eq_rhs)
    if method == "CC2":
        return (eq.T1eq_rhs, eq.T2eq_rhs_CC2, eq.L1eq_rhs_CC2, eq.L2eq_rhs_CC2)
    raise ValueError(method)


def rk4_step(eq, method, F, t1, t2, lam1, lam2, t, dt, Vt, precs):
    """One loop iteration of the reference after `t += dt` has been applied.

    Returns the new state and the rounded base (t1min, t2min, L1min, L2min).
    """
    f_t1, f_t2, f_l1, f_l2 = rhs_set(eq, method)
    L1min = np.around(lam1, decimals=precs)
    L2min = np.around(lam2, decimals=precs)
    t1min = np.around(t1.copy(), decimals=precs)
    t2min = np.around(t2.copy(), decimals=precs)
    dt1 = -1j * chain(lambda y, Fs: f_t1(y, t2, Fs), t1, t, dt, F, Vt)
    dt2 = -1j * chain(lambda y, Fs: f_t2(t1, y, Fs), t2, t, dt, F, Vt)
    dL1 = 1j * chain(lambda y, Fs: f_l1(t1, t2, y, lam2, Fs), lam1, t, dt, F, Vt)
    dL2 = 1j * chain(lambda y, Fs: f_l2(t1, t2, lam1, y, Fs), lam2, t, dt, F, Vt)
    new = (t1min + dt1, t2min + dt2, L1min + dL1, L2min + dL2)
    return new, (t1min, t2min, L1min, L2min)


class Trajectory:
    """What one propagation produced."""

    def __init__(self):
        self.times = []        # time labels as the reference records them
        self.t_exact = []      # unrounded t after each step
        self.mu = []           # (3,) complex dipole after each step
        self.mu0 = None        # dipole before the first step
        self.state = None      # final (t1, t2, lam1, lam2)
        self.saved = None      # (t1min, t2min, L1min, L2min) when a stop rule fired
        self.saved_time = None
        self.stop_reason = None
        self.steps = 0


def propagate(eq, F, t1, t2, lam1, lam2, mu, w0, A, t0, dt, *, method="CCSD",
              pol=2, precs=15, tf=None, max_steps=None, timeout=None,
              observe=None):
    """Run the reference's RK loop.

    `mu` is (3, n, n); the field uses mu[pol] (the reference hard-codes z).
    The loop runs while t < tf (if given) and fewer than max_steps steps were
    taken (if given).  `timeout` is in minutes as in the reference.
    `observe(i, t, state)` is called after each update when given.
    """
    tr = Trajectory()
    Vt = lambda s: field(mu[pol], A, w0, s)
    ndig = 4 if method == "CCSD" else 5
    t = t0
    start = time.time()
    tr.mu0 = np.array(eq.dipole_moment(t1, t2, lam1, lam2, mu))
    i = 0
    while True:
        if tf is not None and not (t < tf):
            break
        if max_steps is not None and i >= max_steps:
            break
        i += 1
        t += dt
        (t1, t2, lam1, lam2), base = rk4_step(eq, method, F, t1, t2, lam1, lam2,
                                              t, dt, Vt, precs)
        mua = np.array(eq.dipole_moment(t1, t2, lam1, lam2, mu))
        tr.times.append(round(t, ndig))
        tr.t_exact.append(t)
        tr.mu.append(mua)
        if observe is not None:
            observe(i, t, (t1, t2, lam1, lam2))
        elapsed = time.time() - start
        if timeout is not None and abs(elapsed) > 0.9 * timeout * 60.0:
            tr.saved, tr.saved_time, tr.stop_reason = base, t - dt, "timeout"
            break
        if abs(mua[2].real) > 100:
            tr.saved, tr.saved_time, tr.stop_reason = base, t - dt, "divergence"
            break
    tr.steps = i
    tr.state = (t1, t2, lam1, lam2)
    tr.mu = np.array(tr.mu).reshape(-1, 3)
    return tr
