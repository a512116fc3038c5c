"""Write golden input/output vectors from the reference implementation itself.

Run in the build container only (the reference tree is not present elsewhere):

    python tests/golden/make_golden.py /root/reference

Why an import is needed: away from the zero-amplitude and fixed-point special
cases, nothing but the reference fixes the exact values of the residuals
(T1eq_rhs, T2eq_rhs, lam_1eq_rhs, lam2eq_rhs, T2eq_rhs_CC2, L1eq_rhs_CC2,
L2eq_rhs_CC2), their intermediates, the density blocks / dipole, or a
Runge_Kutta_solver trajectory under a field.  The CC2 Lambda equations in
particular have no textbook closed form.  This script evaluates exactly those
functions on seeded inputs and stores inputs and outputs as .npz data.

The reference needs psi4 and opt_einsum, neither of which is installed.  Stub
modules stand in for them: `opt_einsum.contract` becomes numpy.einsum with
optimize=True, and the psi4-backed attributes (TEI, occ/vir slices, overlap,
dipole integrals) are set on an object created without running __init__.
Only data leaves this script; no reference source is stored.
"""
from __future__ import annotations

import contextlib
import io
import os
import sys
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from python_code_to_check_c_amd import synthetic  # noqa: E402

CASES = [
    # name, config id, o, v, rk steps
    ("o4v6", 11, 4, 6, 4),
    ("c1_o10v4", 1, 10, 4, 3),
]

T_FUNCS = ["Fae", "Fmi", "Fme", "T1eq_rhs", "Wmnij", "Wmnij_2", "Wabef", "Wabef_2",
           "Wmbej", "T2eq_rhs", "T2eq_rhs_CC2"]
TF_FUNCS = ["LRFea", "LRFim", "LSWieam", "LRWefam", "LRWibjm", "LSWmnij", "LSWabef"]
L_FUNCS = ["lam_1eq_rhs", "lam2eq_rhs", "L1eq_rhs_CC2", "L2eq_rhs_CC2"]
D_FUNCS = ["Dij", "Dab", "Dai"]


def load_reference(ref_dir):
    psi4 = types.ModuleType("psi4")
    sys.modules["psi4"] = psi4
    oe = types.ModuleType("opt_einsum")
    oe.contract = lambda *a, **k: np.einsum(*a, optimize=True)
    sys.modules["opt_einsum"] = oe
    sys.path.insert(0, ref_dir)
    sys.dont_write_bytecode = True  # leave no .pyc inside the read-only reference tree
    import CC2_Helper  # noqa: E402
    return CC2_Helper.CC2_Helper


def rand_amplitudes(rng, o, v, scale):
    def cplx(shape):
        return scale * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
    t1 = cplx((o, v))
    t2 = cplx((o, o, v, v))
    t2 = t2 - t2.transpose(1, 0, 2, 3)
    t2 = t2 - t2.transpose(0, 1, 3, 2)
    return t1, 0.5 * t2


def make_case(cls, name, cid, o, v, nsteps, out_dir):
    n = o + v
    sysm = synthetic.SyntheticSystem(cid, o, v)
    obj = cls.__new__(cls)
    obj.TEI = sysm.TEI
    obj.occ = slice(0, o)
    obj.vir = slice(o, n)
    obj.nmo = n // 2
    obj.ndocc = o // 2
    obj.S = np.eye(n // 2)
    mu = sysm.mu
    obj.Defd_dipole = lambda: [mu[0], mu[1], mu[2]]

    rng = np.random.default_rng(1000 + cid)
    t1, t2 = rand_amplitudes(rng, o, v, 0.05)
    lam1, lam2 = rand_amplitudes(rng, o, v, 0.05)
    A, w0, tt = 0.002, 0.923, 0.0137
    Fc = sysm.F + (-A * mu[2] * np.exp(1j * w0 * 2 * np.pi * tt))

    rec = {"TEI": sysm.TEI, "F": sysm.F, "Fc": Fc, "mu": mu, "t1": t1, "t2": t2,
           "lam1": lam1, "lam2": lam2, "o": o, "v": v}
    for f in T_FUNCS + TF_FUNCS:
        rec["out_" + f] = getattr(obj, f)(t1, t2, Fc)
    rec["out_Gfe"] = obj.Gfe(t2, lam2)
    rec["out_Gmn"] = obj.Gmn(t2, lam2)
    rec["out_LWfiea"] = obj.LWfiea(t1)
    rec["out_LWmina"] = obj.LWmina(t1)
    for f in L_FUNCS:
        rec["out_" + f] = getattr(obj, f)(t1, t2, lam1, lam2, Fc)
    for f in D_FUNCS:
        rec["out_" + f] = getattr(obj, f)(t1, t2, lam1, lam2)
    rec["out_CCSD_Corr_E"] = np.asarray(obj.CCSD_Corr_E(t1, t2, Fc))
    with contextlib.redirect_stdout(io.StringIO()):
        rec["out_dipole_moment"] = np.array(obj.dipole_moment(t1, t2, lam1, lam2, sysm.F))

    # Short Runge_Kutta_solver trajectories: the reference returns nothing and
    # saves nothing when it reaches tf, so the state after each step is taken
    # from the arguments of its per-step dipole_moment call.
    dt = 0.001
    for method, solver in (("CCSD", "Runge_Kutta_solver"), ("CC2", "Runge_Kutta_solver_CC2")):
        seen = []
        orig = obj.dipole_moment

        def spy(a, b, c, d, F, _orig=orig, _seen=seen):
            r = _orig(a, b, c, d, F)
            _seen.append((a.copy(), b.copy(), c.copy(), d.copy(), np.array(r)))
            return r

        obj.dipole_moment = spy
        tf = (nsteps - 0.5) * dt
        with contextlib.redirect_stdout(io.StringIO()):
            getattr(obj, solver)(sysm.F, t1, t2, lam1, lam2, w0, A, 0.0, tf, dt,
                                 1e9, 15)
        obj.dipole_moment = orig
        assert len(seen) == nsteps + 1, len(seen)
        rec[f"rk_{method}_mu"] = np.array([s[4] for s in seen])
        for k, key in enumerate(("t1", "t2", "lam1", "lam2")):
            rec[f"rk_{method}_{key}"] = np.array([s[k] for s in seen[1:]])
    rec["rk_w0"], rec["rk_A"], rec["rk_dt"], rec["rk_steps"] = w0, A, dt, nsteps
    path = os.path.join(out_dir, f"golden_{name}.npz")
    np.savez_compressed(path, **rec)
    print("wrote", path, os.path.getsize(path), "bytes")


def main():
    ref_dir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    cls = load_reference(ref_dir)
    for case in CASES:
        make_case(cls, *case, HERE)


if __name__ == "__main__":
    main()
