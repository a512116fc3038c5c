"""Differentiable windowed dipole of a real-time TDCC propagation.

`windowed_dipole(...)` returns (Re L, Im L) of L = sum_n w_n mu_pol(t_n) over
`nsteps` RK4 steps as a torch tensor that is differentiable with respect to the
field amplitude A, the frequency w0 and the dipole matrix mu_pol (entering the
field -A mu_pol exp(i 2 pi w0 t) and the read-out).  The derivatives come from
the library's hand-written adjoint (tdcc_dipole_gradient: transposed
contractions of the step, checkpoint-and-recompute every `checkpoint_every`
steps), run together with the forward propagation; backward only contracts
them with the incoming gradient.  There is no counterpart in the reference,
which has no derivatives.
"""
from __future__ import annotations

import numpy as np
import torch


class WindowedDipole(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, w0, mu_pol, context, state, method, nsteps, dt, weights, F, mu_all, pol, precs,
                checkpoint_every):
        mu = np.array(mu_all, dtype=float, copy=True)
        mu[pol] = mu_pol.detach().cpu().numpy()
        context.set_system(None, F, mu)
        try:
            context.set_state(*state)
            L, dA, dw, dmu = context.dipole_gradient(method, nsteps, 0.0, dt, float(w0), float(A), weights,
                                                     pol=pol, precs=precs, checkpoint_every=checkpoint_every)
        finally:
            context.set_system(None, F, mu_all)
        ctx.grads = (dA, dw, dmu)
        return torch.tensor([L.real, L.imag], dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        dA, dw, dmu = ctx.grads
        gr, gi = float(g[0]), float(g[1])
        gA = torch.tensor(gr * dA.real + gi * dA.imag, dtype=torch.float64)
        gw = torch.tensor(gr * dw.real + gi * dw.imag, dtype=torch.float64)
        gmu = torch.from_numpy(gr * dmu.real + gi * dmu.imag)
        return (gA, gw, gmu) + (None,) * 11


def windowed_dipole(context, state, method, nsteps, dt, weights, A, w0, mu_pol, F, mu_all, pol=2, precs=15,
                    checkpoint_every=50):
    """(Re L, Im L) with L = sum_n weights[n] mu_pol(t_n), differentiable in
    A, w0 (0-d float64 tensors) and mu_pol ((n, n) float64 tensor).

    `context` is a python_code_to_check_c_amd._lib.Context holding the system
    (F, mu_all = the three dipole matrices); `state` is (t1, t2, lam1, lam2) at
    t = 0.  The context's state is left at the end of the propagation."""
    return WindowedDipole.apply(A, w0, mu_pol, context, state, method, int(nsteps), float(dt),
                                np.asarray(weights, dtype=float), np.asarray(F, dtype=float),
                                np.asarray(mu_all, dtype=float), int(pol), int(precs), int(checkpoint_every))
