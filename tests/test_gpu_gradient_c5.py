"""The windowed-dipole gradient at the differentiable workload's own shape:
o=10, v=38, FP64, 500 RK4 steps at dt = 0.001, L = sum_n hann_n dt mu_z(t_n),
checkpoints every 50 steps with recompute.  The gradients with respect to the
field amplitude A, the frequency w0 and entries of the z dipole matrix are
checked against central finite differences of the GPU propagation itself
(h = 1e-5 relative for A and w0; 1e-5 absolute for dipole entries, which are
O(0.3)).  The state is the seeded complex test state of config id 2 with TEI
sigma 0.02 (python_code_to_check_c_amd.synthetic.perturbed_state).

Tolerances: A and w0 with the reference's rounding to 15 decimals (precs 15,
which the adjoint treats as the identity; its effect on the differences is
below 1e-7 relative here): 2e-6 and 1e-5 relative; dipole entries without
rounding (precs 22): 1e-6 relative.
"""
import numpy as np
import pytest

from python_code_to_check_c_amd import _lib, synthetic

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]
O, V, N, DT, W0, A, CKPT = 10, 38, 500, 0.001, 0.923, 0.002, 50


@pytest.fixture(scope="module")
def c5():
    print("\n[gradient o=10 v=38, 500 steps, checkpoints every 50]", flush=True)
    s = synthetic.SyntheticSystem(2, O, V, tei_sigma=0.02)
    X0 = synthetic.perturbed_state(2, O, V, 0.02)
    ctx = _lib.Context(O, V)
    ctx.set_system(s.TEI, s.F, s.mu)
    w = np.hanning(N + 2)[1:-1] * DT
    try:
        yield s, ctx, X0, w
    finally:
        ctx.close()


def _loss(ctx, s, X0, w, A_=A, w0_=W0, mu_=None, precs=15):
    if mu_ is not None:
        ctx.set_system(None, s.F, mu_)
    try:
        ctx.set_state(*X0)
        mu, _ = ctx.rk4_run("CCSD", N, 0.0, DT, w0_, A_, 2, precs)
    finally:
        if mu_ is not None:
            ctx.set_system(None, s.F, s.mu)
    assert np.all(np.isfinite(mu))
    return np.sum(w * mu[:, 2])


def test_gradient_a_w0_against_finite_differences(c5):
    s, ctx, X0, w = c5
    ctx.set_state(*X0)
    L, dA, dw, _ = ctx.dipole_gradient("CCSD", N, 0.0, DT, W0, A, w, pol=2, precs=15, checkpoint_every=CKPT)
    assert abs(L - _loss(ctx, s, X0, w)) <= 1e-13 * abs(L)
    h = 1e-5 * A
    fdA = (_loss(ctx, s, X0, w, A_=A + h) - _loss(ctx, s, X0, w, A_=A - h)) / (2 * h)
    h = 1e-5 * W0
    fdw = (_loss(ctx, s, X0, w, w0_=W0 + h) - _loss(ctx, s, X0, w, w0_=W0 - h)) / (2 * h)
    print(f"dL/dA {dA:.10e} fd {fdA:.10e} rel {abs(dA - fdA) / abs(fdA):.2e}; "
          f"dL/dw0 {dw:.10e} fd {fdw:.10e} rel {abs(dw - fdw) / abs(fdw):.2e}", flush=True)
    assert abs(dA - fdA) <= 2e-6 * abs(fdA)
    assert abs(dw - fdw) <= 1e-5 * abs(fdw)


def test_gradient_dipole_entries_against_finite_differences(c5):
    s, ctx, X0, w = c5
    ctx.set_state(*X0)
    _, _, _, dmu = ctx.dipole_gradient("CCSD", N, 0.0, DT, W0, A, w, pol=2, precs=22, checkpoint_every=CKPT)
    for p, q in [(2, O + 5), (O + 7, O + 3)]:
        h = 1e-5
        mp, mm = s.mu.copy(), s.mu.copy()
        mp[2, p, q] += h
        mm[2, p, q] -= h
        fd = (_loss(ctx, s, X0, w, mu_=mp, precs=22) - _loss(ctx, s, X0, w, mu_=mm, precs=22)) / (2 * h)
        print(f"dL/dmu[{p},{q}] {dmu[p, q]:.10e} fd {fd:.10e} rel {abs(dmu[p, q] - fd) / abs(fd):.2e}", flush=True)
        assert abs(dmu[p, q] - fd) <= 1e-6 * abs(fd), (p, q)
