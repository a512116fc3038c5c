"""Oracle trajectories at the C2 shape (o=10, v=38) for the 1000-step GPU parity test.

    python tests/golden/make_c2_trajectory.py [steps]

Inputs: the seeded synthetic system (config id 2, TEI sigma 0.02: with the
literal sigma 0.05 the ground-state DIIS does not converge at this shape), the
CCSD (resp. CC2) ground state converged by the oracle's Jacobi + DIIS to
max|residual| < 1e-10, then `steps` RK4 steps at dt = 0.001, z field,
w0 = 0.923, A = 0.002, precs = 15 with the oracle propagator (which matches the
reference's Runge_Kutta_solver, see tests/test_oracle_golden.py).

Stored (compressed): the initial amplitudes, the dipole (x, y, z) after every
step, the exact times, and the final t1 / lam1 (the final doubles are summarized
by their antisymmetry defect).  Antisymmetric (o,o,v,v) arrays are stored as
their i<j, a<b elements.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ground_state, rk4  # noqa: E402
from oracle.cc_equations import CCEquations  # noqa: E402
from python_code_to_check_c_amd import synthetic  # noqa: E402

CONFIG_ID, O, V, SIGMA = 2, 10, 38, 0.02
DT, W0, A = 0.001, 0.923, 0.002


def pack4(x):
    o, v = x.shape[0], x.shape[2]
    i, j = np.triu_indices(o, 1)
    a, b = np.triu_indices(v, 1)
    return x[i[:, None], j[:, None], a[None, :], b[None, :]]


def unpack4(p, o, v):
    i, j = np.triu_indices(o, 1)
    a, b = np.triu_indices(v, 1)
    x = np.zeros((o, o, v, v), dtype=p.dtype)
    x[i[:, None], j[:, None], a[None, :], b[None, :]] = p
    x[j[:, None], i[:, None], a[None, :], b[None, :]] = -p
    x[i[:, None], j[:, None], b[None, :], a[None, :]] = -p
    x[j[:, None], i[:, None], b[None, :], a[None, :]] = p
    return x


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 1000
    s = synthetic.SyntheticSystem(CONFIG_ID, O, V, tei_sigma=SIGMA)
    eq = CCEquations(s.TEI, O)
    rec = {"config_id": CONFIG_ID, "o": O, "v": V, "sigma": SIGMA, "dt": DT, "w0": W0, "A": A, "steps": steps}
    for method in ("CCSD", "CC2"):
        t0 = time.time()
        t1, t2, it = ground_state.solve_t(eq, s.F, method, tol=1e-10, maxiter=300)
        l1, l2, il = ground_state.solve_lambda(eq, s.F, t1, t2, method, tol=1e-10, maxiter=300)
        print(method, "ground state", it, il, f"{time.time() - t0:.1f}s", flush=True)
        for k, x in (("t1", t1), ("l1", l1)):
            rec[f"{method}_{k}0"] = x
        # the converged t2/lam2 are antisymmetric to rounding; keep them exactly so
        t2 = unpack4(pack4(t2), O, V)
        l2 = unpack4(pack4(l2), O, V)
        rec[f"{method}_t2p0"] = pack4(t2)
        rec[f"{method}_l2p0"] = pack4(l2)
        t0 = time.time()
        tr = rk4.propagate(eq, s.F, t1, t2, l1, l2, s.mu, W0, A, 0.0, DT, method=method, max_steps=steps)
        print(method, steps, "steps", f"{time.time() - t0:.1f}s", flush=True)
        rec[f"{method}_mu0"] = tr.mu0
        rec[f"{method}_mu"] = tr.mu
        rec[f"{method}_t"] = np.array(tr.t_exact)
        ft1, ft2, fl1, fl2 = tr.state
        rec[f"{method}_t1"] = ft1
        rec[f"{method}_l1"] = fl1
        rec[f"{method}_t2_asym"] = np.abs(ft2 + ft2.transpose(1, 0, 2, 3)).max()
    path = os.path.join(HERE, "c2_trajectory.npz")
    np.savez_compressed(path, **rec)
    print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    main()
