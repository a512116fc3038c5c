"""Drop-in CC2_Helper (CC2_Helper.py:7-500) for the propagation path, on the HIP library."""
from __future__ import annotations

from . import _lib
from .CCSD_Helper import CCSD_Helper


class CC2_Helper(CCSD_Helper):
    method = "CC2"
    time_digits = 5   # CC2_Helper.py:484 labels time with round(t, 5)

    def T2eq_rhs_CC2(self, t1, t2, F):
        """CC2_Helper.py:12-47"""
        return self.ctx.rhs(_lib.RHS_T2_CC2, F, t1, t2)

    def L1eq_rhs_CC2(self, t1, t2, lam1, lam2, F):
        """CC2_Helper.py:50-125"""
        return self.ctx.rhs(_lib.RHS_L1_CC2, F, t1, t2, lam1, lam2)

    def L2eq_rhs_CC2(self, t1, t2, lam1, lam2, F):
        """CC2_Helper.py:128-175"""
        return self.ctx.rhs(_lib.RHS_L2_CC2, F, t1, t2, lam1, lam2)

    def Runge_Kutta_solver_CC2(self, F, t1, t2, lam1, lam2, w0, A, t0, tf, dt, timeout, precs, restart=None):
        """CC2_Helper.py:437-500 (the CCSD T1 equation with the CC2 T2/Lambda equations)."""
        return self._propagate(F, t1, t2, lam1, lam2, w0, A, t0, tf, dt, timeout, precs, restart)
