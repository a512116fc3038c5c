"""The digit-GEMM autotuner leaves results bit for bit unchanged in its default
mode (TDCC_AUTOTUNE=1: tile orders and persistent grids only, every output
still summed by one block in the same order), so a run's results do not depend
on which candidate happened to time fastest.  Mode 2 also tries fewer K splits,
which reorders FP64 partial sums: within 1e-12 relative of the untuned run.

Workload: o=10, v=38 (config id 2, TEI sigma 0.02, the seeded complex test
state: MP2 amplitudes plus noise, t1 and lam1 nonzero), 5 digits, every
batch-free contraction on the digit-GEMM path (TDCC_GEMM_MIN_FLOPS=0) so that
many launch shapes are tuned; 6 RK4 steps, CCSD.
"""
import os

import numpy as np
import pytest

from python_code_to_check_c_amd import _lib, synthetic

pytestmark = pytest.mark.gpu
O, V, CID, SIGMA = 10, 38, 2, 0.02
STEPS, DT, W0, A = 6, 0.001, 0.923, 0.002
KEYS = ("TDCC_AUTOTUNE", "TDCC_GEMM_MIN_FLOPS")


def run(s, mode):
    saved = {k: os.environ.get(k) for k in KEYS}
    os.environ["TDCC_AUTOTUNE"] = str(mode)
    os.environ["TDCC_GEMM_MIN_FLOPS"] = "0"
    try:
        ctx = _lib.Context(O, V)
        try:
            ctx.set_system(s.TEI, s.F, s.mu)
            ctx.set_precision(5)
            ctx.set_state(*synthetic.perturbed_state(CID, O, V, SIGMA))
            print(f"\n[autotune mode {mode}: o={O} v={V} {STEPS} steps, 5 digits]", flush=True)
            mu, _ = ctx.rk4_run("CCSD", STEPS, 0.0, DT, W0, A)
            return mu
        finally:
            ctx.close()
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.fixture(scope="module")
def system():
    return synthetic.SyntheticSystem(CID, O, V, tei_sigma=SIGMA)


def test_default_autotune_is_bitwise_neutral(system):
    mu0 = run(system, 0)
    mu1 = run(system, 1)
    assert np.all(np.isfinite(mu0))
    assert np.array_equal(mu0, mu1)


def test_k_split_autotune_within_rounding(system):
    mu0 = run(system, 0)
    mu2 = run(system, 2)
    scale = np.abs(mu0).max()
    assert np.abs(mu2 - mu0).max() <= 1e-12 * scale
