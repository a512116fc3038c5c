"""Batched trajectories: one context, several field polarisations / frequencies /
amplitudes, one captured step for all of them.

Two programs: the stacked one (default; one trajectory's step graph with a
trajectory axis, shared integral operands against the slots' stacked rows) and
the side-by-side one (TDCC_STACK=0; each slot's step graph, the same kernels
with the same arguments as a single run).  Side by side, each slot reproduces
the single-trajectory run bit for bit; stacked, it agrees with it to 1e-12
relative (the stacked launches may cut K differently, which changes the order
of FP64 partial sums), and each slot's result does not depend on the other
slots (permuting the slots permutes the results bit for bit).  Both in FP64
and in split precision; one slot is also checked against the oracle (1e-11).
Inputs: the golden o=4, v=6 system.
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


@pytest.fixture(params=["stack", "sides"])
def mode(request, monkeypatch):
    if request.param == "sides":
        monkeypatch.setenv("TDCC_STACK", "0")
    else:
        monkeypatch.delenv("TDCC_STACK", raising=False)
    return request.param


def _close(x, y, exact):
    if exact:
        return np.array_equal(x, y)
    return np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300)


@pytest.mark.parametrize("digits", [0, 4])
@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_batch_equals_single_runs(sysg, method, digits, mode):
    g, ctx = sysg
    print(f"\n[batch {method} digits={digits} {mode}]", flush=True)
    exact = mode == "sides"
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
            assert _close(mu_b[b], mu, exact), b
            single = ctx.get_state(0)
            batched = ctx.get_state_slot(b, 0)
            for x, y in zip(batched, single):
                assert _close(x, y, exact), b
            base_s, base_b = ctx.get_state(1), ctx.get_state_slot(b, 1)
            for x, y in zip(base_b, base_s):
                assert _close(x, y, exact), b
    finally:
        ctx.set_batch(0)
        ctx.set_precision(0)


@pytest.mark.parametrize("digits", [0, 4])
def test_stacked_slots_are_independent(sysg, digits, monkeypatch):
    """Reversing the slots (states and field parameters) reverses the stacked
    results bit for bit: no slot's arithmetic depends on another's values."""
    monkeypatch.delenv("TDCC_STACK", raising=False)
    g, ctx = sysg
    ctx.set_precision(digits)
    try:
        nb = len(POL)
        out = []
        for order in (list(range(nb)), list(range(nb))[::-1]):
            ctx.set_batch(nb)
            for k, b in enumerate(order):
                ctx.set_state_slot(k, *slot_state(g, b))
            mu, _ = ctx.rk4_run_batch("CCSD", STEPS, 0.0, DT, [POL[b] for b in order], [W0[b] for b in order],
                                      [AMP[b] for b in order])
            out.append({b: (mu[k], ctx.get_state_slot(k, 0)) for k, b in enumerate(order)})
        for b in range(nb):
            assert np.array_equal(out[0][b][0], out[1][b][0]), b
            for x, y in zip(out[0][b][1], out[1][b][1]):
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
