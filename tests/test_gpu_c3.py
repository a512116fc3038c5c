"""The HIP path at o=20, v=160 (the split-precision workload's shape) against
the oracle's sampled residuals (tests/golden/c3_samples.npz, written by
tests/golden/make_c3_samples.py).

At this shape the default lowering takes the routes that smaller shapes never
reach: the banded digit-GEMM tile order, K-split ladders with their partial
sums, the three-pass digit split of long rows, the persistent digit GEMM, the
column-tile split kernels above 2048 rows, the FP64 dense GEMM on 300 ladder
tiles and the multi-tile dot kernel.  Inputs: the synthetic system of config id
3 with TEI sigma 0.05 (the TEI generated on the device), the seeded complex
state synthetic.perturbed_state(3, 20, 160) and F + V(t) at t = 0.0137.

Bars (relative to the residual's largest element, stored with the samples):
FP64 1e-12.  Split precision with S digits: 300 * 128^-S, ten times the
constant of the smaller-shape split tests, since the sums here run over up to
v^2 = 25,600 terms and the residuals at this state are small differences of
much larger terms (T2eq_rhs_CC2 at S = 4: 1.5e-7 of its largest element, 1.8e-8
absolute); and one more digit must shrink every residual's error by at least
32 (the truncation is 128^-S per operand element), which a dropped or misplaced
digit product would not do.  (The digits are radix 254, so the truncation is
0.5 * 254^-S of the row scale and these bars, set for radix 128, hold with
room to spare.)
"""
import os

import numpy as np
import pytest

from python_code_to_check_c_amd import _lib, synthetic

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
RHS = [("T1eq_rhs", _lib.RHS_T1), ("T2eq_rhs", _lib.RHS_T2), ("lam_1eq_rhs", _lib.RHS_L1),
       ("lam2eq_rhs", _lib.RHS_L2), ("T2eq_rhs_CC2", _lib.RHS_T2_CC2), ("L1eq_rhs_CC2", _lib.RHS_L1_CC2),
       ("L2eq_rhs_CC2", _lib.RHS_L2_CC2)]


@pytest.fixture(scope="module")
def c3():
    from oracle import rk4
    g = dict(np.load(os.path.join(HERE, "golden", "c3_samples.npz")))
    o, v, cid, sigma = int(g["o"]), int(g["v"]), int(g["config_id"]), float(g["sigma"])
    print(f"\n[c3 case o={o} v={v} config {cid} sigma {sigma}]", flush=True)
    s = synthetic.SyntheticSystem(cid, o, v, tei_sigma=sigma, tei_on_device=True)
    state = synthetic.perturbed_state(cid, o, v, sigma)
    Fc = s.F + rk4.field(s.mu[2], float(g["A"]), float(g["w0"]), float(g["t_field"]))
    ctx = _lib.Context(o, v)
    try:
        ctx.set_system(None, s.F, s.mu)
        ctx.synth_tei(cid, sigma)
        yield g, ctx, state, Fc
    finally:
        ctx.close()


def _rel_errors(g, ctx, state, Fc):
    out = {}
    for name, kind in RHS:
        got = ctx.rhs(kind, Fc, *state)
        if got.ndim == 4:
            got = got.reshape(-1)[g["index"]]
        out[name] = float(np.abs(got - g["out_" + name]).max() / float(g["maxabs_" + name]))
    return out


def test_fp64_residuals(c3):
    g, ctx, state, Fc = c3
    ctx.set_precision(0)
    rel = _rel_errors(g, ctx, state, Fc)
    print("FP64 relative errors", rel, flush=True)
    for name, r in rel.items():
        assert r <= 1e-12, (name, r)


def test_fp64_dipole(c3):
    g, ctx, state, Fc = c3
    ctx.set_precision(0)
    mu = ctx.dipole(*state)
    ref = g["out_dipole_moment"]
    assert np.abs(mu - ref).max() <= 1e-12 * np.abs(ref).max()


def test_split_residuals(c3):
    g, ctx, state, Fc = c3
    rel = {}
    try:
        for S in (4, 5):
            ctx.set_precision(S)
            rel[S] = _rel_errors(g, ctx, state, Fc)
            print(f"split S={S} relative errors", rel[S], flush=True)
    finally:
        ctx.set_precision(0)
    for S, errs in rel.items():
        for name, r in errs.items():
            assert r <= 300.0 * 128.0 ** -S, (S, name, r)
    for name in rel[4]:
        assert rel[5][name] * 32.0 <= rel[4][name], (name, rel[4][name], rel[5][name])


@pytest.mark.parametrize("digits", [0, 5])
def test_two_stream_steps_at_workload_shape(digits):
    """Two RK4 steps of CCSD at o=20, v=160 from the seeded state: the step
    with its T side and Lambda side on two streams (the default) gives
    bitwise the dipoles and amplitudes of the one-stream step
    (TDCC_STREAMS=1), in FP64 and with 5 digits.  At this shape the step holds
    the packed ladders, the three-product rings and the T-side buffers the
    Lambda side reads across the streams."""
    g = dict(np.load(os.path.join(HERE, "golden", "c3_samples.npz")))
    o, v, cid, sigma = int(g["o"]), int(g["v"]), int(g["config_id"]), float(g["sigma"])
    s = synthetic.SyntheticSystem(cid, o, v, tei_sigma=sigma, tei_on_device=True)
    state = synthetic.perturbed_state(cid, o, v, sigma)
    out = {}
    saved = os.environ.get("TDCC_STREAMS")
    try:
        for mode in ("2", "1"):
            os.environ["TDCC_STREAMS"] = mode
            print(f"\n[c3 steps digits={digits} streams={mode}]", flush=True)
            ctx = _lib.Context(o, v)
            try:
                ctx.set_system(None, s.F, s.mu)
                ctx.synth_tei(cid, sigma)
                ctx.set_precision(digits)
                ctx.set_state(*state)
                mu, _ = ctx.rk4_run("CCSD", 2, 0.0, 0.001, 0.923, 0.002)
                out[mode] = [mu, *ctx.get_state(0)]
            finally:
                ctx.close()
    finally:
        if saved is None:
            os.environ.pop("TDCC_STREAMS", None)
        else:
            os.environ["TDCC_STREAMS"] = saved
    assert np.isfinite(out["2"][0]).all()
    for x, y in zip(out["2"], out["1"]):
        assert np.array_equal(x, y)
