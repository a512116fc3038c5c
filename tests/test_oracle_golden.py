"""The oracle against golden vectors written from the reference itself.

tests/golden/golden_*.npz hold inputs and the reference's outputs for every
residual, intermediate, density block, the dipole and short Runge_Kutta_solver
trajectories (CCSD and CC2); see tests/golden/make_golden.py.
"""
import os

import numpy as np
import pytest

from oracle import rk4
from oracle.cc_equations import CCEquations

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["o4v6", "c1_o10v4"]
TOL = 1e-13

T_F = ["Fae", "Fmi", "Fme", "T1eq_rhs", "Wmnij", "Wmnij_2", "Wabef", "Wabef_2", "Wmbej", "T2eq_rhs",
       "T2eq_rhs_CC2", "LRFea", "LRFim", "LSWieam", "LRWefam", "LRWibjm", "LSWmnij", "LSWabef"]
L_F = ["lam_1eq_rhs", "lam2eq_rhs", "L1eq_rhs_CC2", "L2eq_rhs_CC2"]
D_F = ["Dij", "Dab", "Dai"]


def load(name):
    g = dict(np.load(os.path.join(HERE, "golden", f"golden_{name}.npz")))
    return g, CCEquations(g["TEI"], int(g["o"]))


@pytest.mark.parametrize("name", CASES)
def test_functions(name):
    g, eq = load(name)
    t1, t2, l1, l2, Fc = g["t1"], g["t2"], g["lam1"], g["lam2"], g["Fc"]
    for f in T_F:
        assert np.abs(getattr(eq, f)(t1, t2, Fc) - g["out_" + f]).max() <= TOL, f
    for f in L_F:
        assert np.abs(getattr(eq, f)(t1, t2, l1, l2, Fc) - g["out_" + f]).max() <= TOL, f
    for f in D_F:
        assert np.abs(getattr(eq, f)(t1, t2, l1, l2) - g["out_" + f]).max() <= TOL, f
    assert np.abs(eq.Gfe(t2, l2) - g["out_Gfe"]).max() <= TOL
    assert np.abs(eq.Gmn(t2, l2) - g["out_Gmn"]).max() <= TOL
    assert np.abs(eq.LWfiea(t1) - g["out_LWfiea"]).max() <= TOL
    assert np.abs(eq.LWmina(t1) - g["out_LWmina"]).max() <= TOL
    assert abs(eq.CCSD_Corr_E(t1, t2, Fc) - g["out_CCSD_Corr_E"]) <= TOL
    mu = np.array(eq.dipole_moment(t1, t2, l1, l2, g["mu"]))
    assert np.abs(mu - g["out_dipole_moment"]).max() <= TOL


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_rk4_trajectory(name, method):
    g, eq = load(name)
    tr = rk4.propagate(eq, g["F"], g["t1"], g["t2"], g["lam1"], g["lam2"], g["mu"], float(g["rk_w0"]),
                       float(g["rk_A"]), 0.0, float(g["rk_dt"]), method=method, max_steps=int(g["rk_steps"]))
    ref_mu = g[f"rk_{method}_mu"]
    assert np.abs(tr.mu0 - ref_mu[0]).max() <= TOL
    assert np.abs(tr.mu - ref_mu[1:]).max() <= TOL
    for k, key in enumerate(("t1", "t2", "lam1", "lam2")):
        assert np.abs(tr.state[k] - g[f"rk_{method}_{key}"][-1]).max() <= TOL, key


def test_c2_trajectory_fixture_matches_oracle():
    """The first steps of the committed o=10, v=38 trajectories are reproduced
    by the oracle from the stored initial amplitudes (the full 1000 steps take
    about an hour on the CPU; tests/golden/make_c2_trajectory.py)."""
    from python_code_to_check_c_amd import synthetic
    from oracle import rk4
    g = dict(np.load(os.path.join(HERE, "golden", "c2_trajectory.npz")))
    o, v = int(g["o"]), int(g["v"])
    s = synthetic.SyntheticSystem(int(g["config_id"]), o, v, tei_sigma=float(g["sigma"]))
    eq = CCEquations(s.TEI, o)
    i, j = np.triu_indices(o, 1)
    a, b = np.triu_indices(v, 1)

    def unpack(p):
        x = np.zeros((o, o, v, v), complex)
        x[i[:, None], j[:, None], a[None, :], b[None, :]] = p
        x[j[:, None], i[:, None], a[None, :], b[None, :]] = -p
        x[i[:, None], j[:, None], b[None, :], a[None, :]] = -p
        x[j[:, None], i[:, None], b[None, :], a[None, :]] = p
        return x

    for method in ("CCSD", "CC2"):
        tr = rk4.propagate(eq, s.F, g[f"{method}_t10"], unpack(g[f"{method}_t2p0"]), g[f"{method}_l10"],
                           unpack(g[f"{method}_l2p0"]), s.mu, float(g["w0"]), float(g["A"]), 0.0,
                           float(g["dt"]), method=method, max_steps=2)
        assert np.abs(np.array(tr.mu) - g[f"{method}_mu"][:2]).max() <= 1e-13


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("vchunk", [1, 4])
def test_chunked_v4_matches_golden(name, vchunk):
    """The slab-by-slab evaluation of the v^4 intermediates (used at o=20,
    v=160, where they do not fit in host memory whole) reproduces the
    reference's own vectors for every residual that reads them."""
    g = dict(np.load(os.path.join(HERE, "golden", f"golden_{name}.npz")))
    eq = CCEquations(g["TEI"], int(g["o"]), vchunk=vchunk)
    t1, t2, l1, l2, Fc = g["t1"], g["t2"], g["lam1"], g["lam2"], g["Fc"]
    for f in ("T2eq_rhs", "T2eq_rhs_CC2", "LRWefam"):
        assert np.abs(getattr(eq, f)(t1, t2, Fc) - g["out_" + f]).max() <= TOL, f
    for f in ("lam_1eq_rhs", "lam2eq_rhs", "L1eq_rhs_CC2"):
        assert np.abs(getattr(eq, f)(t1, t2, l1, l2, Fc) - g["out_" + f]).max() <= TOL, f
    v = int(g["v"])
    for f in ("Wabef", "Wabef_2", "LSWabef"):
        full = g["out_" + f]
        for a0 in range(0, v, vchunk):
            rows = slice(a0, min(v, a0 + vchunk))
            assert np.abs(getattr(eq, f)(t1, t2, Fc, rows) - full[rows]).max() <= TOL, f


def test_chunked_v4_matches_whole_evaluation():
    """At a shape with uneven slabs (v = 22, slabs of 5) the chunked and the
    whole evaluations agree for every residual."""
    from python_code_to_check_c_amd import synthetic
    o, v = 6, 22
    s = synthetic.SyntheticSystem(5, o, v)
    rng = np.random.default_rng(3)

    def amp(shape):
        x = 0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        if len(shape) == 4:
            x = x - x.transpose(1, 0, 2, 3)
            x = x - x.transpose(0, 1, 3, 2)
        return x

    t1, t2, l1, l2 = amp((o, v)), amp((o, o, v, v)), amp((o, v)), amp((o, o, v, v))
    Fc = s.F + 0.01j * s.mu[2]
    whole, chunked = CCEquations(s.TEI, o), CCEquations(s.TEI, o, vchunk=5)
    for f in ("T2eq_rhs", "T2eq_rhs_CC2"):
        x, y = getattr(whole, f)(t1, t2, Fc), getattr(chunked, f)(t1, t2, Fc)
        assert np.abs(x - y).max() <= 1e-13 * np.abs(x).max(), f
    for f in ("lam_1eq_rhs", "lam2eq_rhs", "L1eq_rhs_CC2"):
        x, y = getattr(whole, f)(t1, t2, l1, l2, Fc), getattr(chunked, f)(t1, t2, l1, l2, Fc)
        assert np.abs(x - y).max() <= 1e-13 * np.abs(x).max(), f
