"""Drop-in CC_Calculator (CC_Calculator.py:28-267) driving the HIP propagation path.

`psi` is a system provider (see CCSD_Helper); the keyword arguments, the
method dispatch, the hard-coded tf = 10.1 and the restart reader follow the
reference.  Missing keyword arguments are ignored as in the reference
(CC_Calculator.py:44-45).
"""
from __future__ import annotations

from . import ground_state, restart_io
from .CC2_Helper import CC2_Helper
from .CCSD_Helper import CCSD_Helper


class CC_Calculator(object):

    def __init__(self, psi, **kwargs):
        self.psi = psi
        self.tol = kwargs.pop("tol", 1e-10)
        self.digits = int(kwargs.pop("digits", 0))  # split-precision propagation (0: FP64)
        self.mol = CCSD_Helper(psi, digits=self.digits)
        self.ndocc = self.mol.ndocc
        try:
            self.w0 = kwargs['w0']
            self.A = kwargs['A']
            self.t0 = kwargs['t0']
            self.tf = 10.1
            self.dt = kwargs['dt']
            self.precs = kwargs['precs']
        except Exception:
            pass

    def TDCC(self, timeout, CCSD_or_CC2):
        """CC_Calculator.py:56-63"""
        if CCSD_or_CC2 == 'CC2':
            self.TDCC2(timeout)
        elif CCSD_or_CC2 == 'CCSD':
            self.TDCCSD(timeout)
        else:
            print("Error in specifying whether it is a CCSD or CC2 calculation")
            print("Correct format is:\nmol.TDCC(stop_time, CCSD)\nwhere CCSD='CCSD' or 'CC2'")

    def TDCC_restart(self, timeout, exact=False):
        """CC_Calculator.py:70-145: resume from the files a stopped run saved.

        exact=True (no reference counterpart) resumes from the binary snapshot
        instead: the state after the stopping step at its exact time, so the
        resumed run continues the interrupted one bit for bit; tf is the
        reference's 50 + the saved t0."""
        if exact:
            r = restart_io.read_state()
            r["tf"] = restart_io.read_parameters()["tf"]
            r["t0"] = r["t"]
        else:
            r = restart_io.read_restart()
        if r["method"] == 'CC2':
            mol = CC2_Helper(self.psi, digits=self.digits)
            mol.Runge_Kutta_solver_CC2(r["F"], r["t1"], r["t2"], r["lam1"], r["lam2"], r["w0"], r["A"],
                                       r["t0"], r["tf"], r["dt"], timeout, r["precs"], 'restart')
        elif r["method"] == 'CCSD':
            mol = self.mol
            mol.Runge_Kutta_solver(r["F"], r["t1"], r["t2"], r["lam1"], r["lam2"], r["w0"], r["A"],
                                   r["t0"], r["tf"], r["dt"], timeout, r["precs"], 'restart')
        self.last = mol

    def _ground_state(self, mol):
        # the ground state is always converged in FP64; only the propagation
        # uses the split precision
        mol.ctx.set_precision(0)
        try:
            F = mol.F_MO()
            t1, t2, _ = ground_state.solve_t(mol, F, tol=self.tol, verbose=True)
            print("E_corr =", mol.CCSD_Corr_E(t1, t2, F))
            lam1, lam2, _ = ground_state.solve_lambda(mol, F, t1, t2, tol=self.tol, verbose=True)
        finally:
            mol.ctx.set_precision(mol.digits)
        return F, t1, t2, lam1, lam2

    def TDCCSD(self, timeout):
        """CC_Calculator.py:155-204"""
        mol = self.mol
        F, t1, t2, lam1, lam2 = self._ground_state(mol)
        mol.Runge_Kutta_solver(F, t1, t2, lam1, lam2, self.w0, self.A, self.t0, self.tf, self.dt,
                               timeout, self.precs)
        self.last = mol

    def TDCC2(self, timeout):
        """CC_Calculator.py:223-267"""
        mol = CC2_Helper(self.psi, digits=self.digits)
        F, t1, t2, lam1, lam2 = self._ground_state(mol)
        mol.Runge_Kutta_solver_CC2(F, t1, t2, lam1, lam2, self.w0, self.A, self.t0, self.tf, self.dt,
                                   timeout, self.precs)
        self.last = mol
