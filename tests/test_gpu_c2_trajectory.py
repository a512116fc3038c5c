"""1000-step TDCCSD / TDCC2 trajectories at o=10, v=38 against the oracle.

tests/golden/c2_trajectory.npz (written by tests/golden/make_c2_trajectory.py)
holds the oracle's ground state (Jacobi + DIIS to |r| < 1e-10) and its dipole
after each of 1000 RK4 steps (dt = 0.001, z field, w0 = 0.923, A = 0.002,
precs = 15), plus the final t1 / lam1.  The GPU path starts from the same
amplitudes on the same seeded system (config id 2, TEI sigma 0.02).

Tolerances: FP64 dipole within 1e-9 a.u. at every step and final t1 / lam1
within 1e-10; split precision (4 int8 digits) dipole within 1e-8 a.u.
"""
import os

import numpy as np
import pytest

from python_code_to_check_c_amd import _lib, synthetic

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "c2_trajectory.npz")


def unpack4(p, o, v):
    i, j = np.triu_indices(o, 1)
    a, b = np.triu_indices(v, 1)
    x = np.zeros((o, o, v, v), dtype=p.dtype)
    x[i[:, None], j[:, None], a[None, :], b[None, :]] = p
    x[j[:, None], i[:, None], a[None, :], b[None, :]] = -p
    x[i[:, None], j[:, None], b[None, :], a[None, :]] = -p
    x[j[:, None], i[:, None], b[None, :], a[None, :]] = p
    return x


@pytest.fixture(scope="module")
def c2():
    g = dict(np.load(FIXTURE))
    o, v = int(g["o"]), int(g["v"])
    s = synthetic.SyntheticSystem(int(g["config_id"]), o, v, tei_sigma=float(g["sigma"]))
    ctx = _lib.Context(o, v)
    ctx.set_system(s.TEI, s.F, s.mu)
    try:
        yield g, ctx, o, v
    finally:
        ctx.close()


def run(g, ctx, o, v, method):
    ctx.set_state(g[f"{method}_t10"], unpack4(g[f"{method}_t2p0"], o, v), g[f"{method}_l10"],
                  unpack4(g[f"{method}_l2p0"], o, v))
    n = int(g["steps"])
    print(f"\n[c2 {method} {n} steps digits={getattr(ctx, 'digits', 0)}]", flush=True)
    mu, ts = ctx.rk4_run(method, n, 0.0, float(g["dt"]), float(g["w0"]), float(g["A"]))
    return mu, ts


@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_fp64_1000_steps(c2, method):
    g, ctx, o, v = c2
    ctx.set_precision(0)
    mu, ts = run(g, ctx, o, v, method)
    assert np.array_equal(ts, g[f"{method}_t"])
    assert np.abs(mu - g[f"{method}_mu"]).max() <= 1e-9
    t1, t2, l1, l2 = ctx.get_state(0)
    assert np.abs(t1 - g[f"{method}_t1"]).max() <= 1e-10
    assert np.abs(l1 - g[f"{method}_l1"]).max() <= 1e-10
    # the propagated doubles stay antisymmetric as in the oracle
    assert np.abs(t2 + t2.transpose(1, 0, 2, 3)).max() <= max(1e-14, 10 * float(g[f"{method}_t2_asym"]))


@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_split_1000_steps(c2, method):
    g, ctx, o, v = c2
    ctx.set_precision(4)
    try:
        mu, _ = run(g, ctx, o, v, method)
    finally:
        ctx.set_precision(0)
    assert np.abs(mu - g[f"{method}_mu"]).max() <= 1e-8
