"""Batched trajectories: one context, several field polarisations / frequencies /
amplitudes, one captured step for all of them.

Each slot must reproduce, bit for bit, the single-trajectory run with the same
parameters (the batched program runs the same kernels with the same
arguments), in FP64 and in split precision; one slot is also checked against
the oracle (1e-11).  Inputs: the golden o=4, v=6 system.
"""
import os

import numpy as np
import pytest

from oracle import rk4
from oracle.cc_equations import CCEquations
from python_code_to_check_c_amd import _lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
POL = [0, 1, 2, 2]
W0 = [0.923, 0.5, 1.1, 0.923]
AMP = [0.002, 0.003, 0.001, 0.002]
DT, STEPS = 0.001, 6


@pytest.fixture(scope="module")
def sysg():
    g = dict(np.load(os.path.join(HERE, "golden", "golden_o4v6.npz")))
    ctx = _lib.Context(int(g["o"]), int(g["v"]))
    ctx.set_system(g["TEI"], g["F"], g["mu"])
    try:
        yield g, ctx
    finally:
        ctx.close()


def slot_state(g, b):
    f = 1.0 + 0.1 * b
    return g["t1"] * f, g["t2"] * f, g["lam1"] * f, g["lam2"] * f


@pytest.mark.parametrize("digits", [0, 4])
@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_batch_equals_single_runs(sysg, method, digits):
    g, ctx = sysg
    print(f"\n[batch {method} digits={digits}]", flush=True)
    ctx.set_precision(digits)
    try:
        nb = len(POL)
        ctx.set_batch(nb)
        for b in range(nb):
            ctx.set_state_slot(b, *slot_state(g, b))
        mu_b, ts_b = ctx.rk4_run_batch(method, STEPS, 0.0, DT, POL, W0, AMP)
        for b in range(nb):
            ctx.set_state(*slot_state(g, b))
            mu, ts = ctx.rk4_run(method, STEPS, 0.0, DT, W0[b], AMP[b], POL[b])
            assert np.array_equal(ts, ts_b)
            assert np.array_equal(mu, mu_b[b]), b
            single = ctx.get_state(0)
            batched = ctx.get_state_slot(b, 0)
            for x, y in zip(single, batched):
                assert np.array_equal(x, y), b
            base_s, base_b = ctx.get_state(1), ctx.get_state_slot(b, 1)
            for x, y in zip(base_s, base_b):
                assert np.array_equal(x, y), b
    finally:
        ctx.set_batch(0)
        ctx.set_precision(0)


def test_batch_slot_against_oracle(sysg):
    g, ctx = sysg
    nb = len(POL)
    ctx.set_batch(nb)
    try:
        for b in range(nb):
            ctx.set_state_slot(b, *slot_state(g, b))
        mu_b, _ = ctx.rk4_run_batch("CCSD", STEPS, 0.0, DT, POL, W0, AMP)
    finally:
        ctx.set_batch(0)
    eq = CCEquations(g["TEI"], int(g["o"]))
    for b in (0, 2):
        tr = rk4.propagate(eq, g["F"], *slot_state(g, b), g["mu"], W0[b], AMP[b], 0.0, DT, pol=POL[b],
                           max_steps=STEPS)
        assert np.abs(mu_b[b] - np.array(tr.mu)).max() <= 1e-11


def test_batch_errors(sysg):
    _, ctx = sysg
    ctx.set_batch(0)
    with pytest.raises(_lib.TdccError):
        ctx.rk4_run_batch("CCSD", 1, 0.0, DT, [], [], [])
    ctx.set_batch(2)
    try:
        with pytest.raises(_lib.TdccError):
            ctx.rk4_run_batch("CCSD", 1, 0.0, DT, [0, 3], [0.9, 0.9], [0.002, 0.002])
        with pytest.raises(_lib.TdccError):
            ctx.get_state_slot(2)
    finally:
        ctx.set_batch(0)
