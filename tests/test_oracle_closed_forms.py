"""Oracle pins that need no reference: closed forms, fixed points, invariants."""
import numpy as np
import pytest

from oracle.cc_equations import CCEquations
from python_code_to_check_c_amd import synthetic

O, V = 4, 6


@pytest.fixture(scope="module")
def sysm():
    s = synthetic.SyntheticSystem(21, O, V)
    return s, CCEquations(s.TEI, O)


def rand_amp(rng, scale=0.05):
    def c(shape):
        return scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    t1 = c((O, V))
    t2 = c((O, O, V, V))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    return t1, t2


def test_zero_amplitudes(sysm):
    """T1eq_rhs(0,0,F) = F_ov; T2eq_rhs(0,0,F) = <ij||ab>; the Lambda residuals
    at t = lam = 0 are F_ov and <ij||ab> (CCSD_Helper.py:186, :281, :572, :625)."""
    s, eq = sysm
    z1, z2 = np.zeros((O, V), complex), np.zeros((O, O, V, V), complex)
    Fov, oovv = s.F[:O, O:], s.TEI[:O, :O, O:, O:]
    assert np.array_equal(eq.T1eq_rhs(z1, z2, s.F), Fov)
    assert np.array_equal(eq.T2eq_rhs(z1, z2, s.F), oovv)
    assert np.array_equal(eq.lam_1eq_rhs(z1, z2, z1, z2, s.F), Fov)
    assert np.array_equal(eq.lam2eq_rhs(z1, z2, z1, z2, s.F), oovv)
    assert np.array_equal(eq.T2eq_rhs_CC2(z1, z2, s.F), oovv)
    assert np.array_equal(eq.L1eq_rhs_CC2(z1, z2, z1, z2, s.F), Fov)
    assert np.array_equal(eq.L2eq_rhs_CC2(z1, z2, z1, z2, s.F), oovv)


def test_cc2_mp2_fixed_point(sysm):
    """With t1 = 0 and diagonal F the CC2 T2 residual is <ij||ab> + (f_aa+f_bb-f_ii-f_jj) t_ijab,
    whose zero is the MP2 amplitude of MP2_E (CC2_Helper.py:22-30, CCSD_Helper.py:111-115)."""
    s, eq = sysm
    F = np.diag(np.diag(s.F))
    e = np.diag(F)
    eo, ev = e[:O], e[O:]
    D = eo.reshape(-1, 1, 1, 1) + eo.reshape(-1, 1, 1) - ev.reshape(-1, 1) - ev
    oovv = s.TEI[:O, :O, O:, O:]
    rng = np.random.default_rng(5)
    _, t2 = rand_amp(rng)
    r = eq.T2eq_rhs_CC2(np.zeros((O, V)), t2, F)
    assert np.abs(r - (oovv - D * t2)).max() < 1e-14
    t2mp2 = oovv / D
    assert np.abs(eq.T2eq_rhs_CC2(np.zeros((O, V)), t2mp2, F)).max() < 1e-14


def test_antisymmetry_preserved(sysm):
    s, eq = sysm
    rng = np.random.default_rng(7)
    t1, t2 = rand_amp(rng)
    l1, l2 = rand_amp(rng)
    F = s.F + 0.01j * s.mu[2]
    for r in (eq.T2eq_rhs(t1, t2, F), eq.lam2eq_rhs(t1, t2, l1, l2, F), eq.T2eq_rhs_CC2(t1, t2, F),
              eq.L2eq_rhs_CC2(t1, t2, l1, l2, F)):
        assert np.abs(r + r.transpose(1, 0, 2, 3)).max() < 1e-13
        assert np.abs(r + r.transpose(0, 1, 3, 2)).max() < 1e-13


def test_lambda_residuals_affine_in_lambda(sysm):
    """At fixed t the Lambda residuals are affine in (lam1, lam2)."""
    s, eq = sysm
    rng = np.random.default_rng(9)
    t1, t2 = rand_amp(rng)
    a1, a2 = rand_amp(rng)
    b1, b2 = rand_amp(rng)
    F = s.F + 0.01j * s.mu[2]
    x, y = 0.3 - 0.2j, 0.7 + 0.1j
    for f in (eq.lam_1eq_rhs, eq.lam2eq_rhs, eq.L1eq_rhs_CC2, eq.L2eq_rhs_CC2):
        z0 = f(t1, t2, 0 * a1, 0 * a2, F)
        lhs = f(t1, t2, x * a1 + y * b1, x * a2 + y * b2, F) - z0
        rhs = x * (f(t1, t2, a1, a2, F) - z0) + y * (f(t1, t2, b1, b2, F) - z0)
        assert np.abs(lhs - rhs).max() < 1e-13


def test_holomorphic(sysm):
    """No conjugation anywhere: with real integrals and F, R(conj x) = conj R(x)."""
    s, eq = sysm
    rng = np.random.default_rng(11)
    t1, t2 = rand_amp(rng)
    l1, l2 = rand_amp(rng)
    c = np.conj
    assert np.abs(eq.T2eq_rhs(c(t1), c(t2), s.F) - c(eq.T2eq_rhs(t1, t2, s.F))).max() < 1e-14
    assert np.abs(eq.lam_1eq_rhs(c(t1), c(t2), c(l1), c(l2), s.F) - c(eq.lam_1eq_rhs(t1, t2, l1, l2, s.F))).max() < 1e-14
