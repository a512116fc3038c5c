"""Ground state on the GPU residuals (python_code_to_check_c_amd.ground_state:
Jacobi + DIIS, the reference's DIIS_solver / DIIS_solver_Lam, CCSD_Helper.py:
347-416, :702-776, and CC2_Helper.py:179-248, :324-400) against the oracle's
solver on the seeded o=10, v=4 system (the small workload): the converged
amplitudes agree to 1e-9, and the GPU residuals at them are below the 1e-10
convergence threshold."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import ground_state as og
from oracle.cc_equations import CCEquations
from python_code_to_check_c_amd import ground_state, synthetic
from python_code_to_check_c_amd.CC2_Helper import CC2_Helper
from python_code_to_check_c_amd.CCSD_Helper import CCSD_Helper

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    s = synthetic.SyntheticSystem(1, 10, 4, tei_sigma=0.05)
    return s, SimpleNamespace(nocc=10, nvir=4, F=s.F, mu=s.mu, TEI=s.TEI)


@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_ground_state_matches_oracle(system, method):
    s, prov = system
    print(f"\n[ground state {method}]", flush=True)
    mol = (CCSD_Helper if method == "CCSD" else CC2_Helper)(prov)
    mol.verbose = False
    t1, t2, _ = ground_state.solve_t(mol, s.F, tol=1e-10)
    l1, l2, _ = ground_state.solve_lambda(mol, s.F, t1, t2, tol=1e-10)
    eq = CCEquations(s.TEI, 10)
    r1, r2, _ = og.solve_t(eq, s.F, method=method, tol=1e-10)
    m1, m2, _ = og.solve_lambda(eq, s.F, r1, r2, method=method, tol=1e-10)
    for got, ref in ((t1, r1), (t2, r2), (l1, m1), (l2, m2)):
        assert np.abs(got - ref).max() <= 1e-9
    Fc = s.F.astype(complex)
    t2rhs = mol.T2eq_rhs if method == "CCSD" else mol.T2eq_rhs_CC2
    assert np.abs(mol.T1eq_rhs(t1, t2, Fc)).max() < 1e-10
    assert np.abs(t2rhs(t1, t2, Fc)).max() < 1e-10
    assert abs(mol.CCSD_Corr_E(t1, t2, s.F) - eq.CCSD_Corr_E(r1, r2, s.F)) <= 1e-12
