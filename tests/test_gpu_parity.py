"""GPU parity of the HIP path against golden vectors written from the reference.

Each case loads the inputs stored in tests/golden/golden_*.npz, runs them
through the C ABI, and compares with the reference's own outputs.
Tolerance: 1e-11 max-abs (FP64 throughout; values are O(1)).
"""
import os

import numpy as np
import pytest

from python_code_to_check_c_amd import _lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-11

INTERMEDIATES = ["Fae", "Fmi", "Fme", "Wmnij", "Wmnij_2", "Wabef", "Wabef_2", "Wmbej", "LRFea", "LRFim",
                 "LSWieam", "LRWefam", "LRWibjm", "LSWmnij", "LSWabef", "Gfe", "Gmn", "LWfiea", "LWmina",
                 "Dij", "Dab", "Dai"]
RHS = [("T1eq_rhs", _lib.RHS_T1), ("T2eq_rhs", _lib.RHS_T2), ("lam_1eq_rhs", _lib.RHS_L1),
       ("lam2eq_rhs", _lib.RHS_L2), ("T2eq_rhs_CC2", _lib.RHS_T2_CC2), ("L1eq_rhs_CC2", _lib.RHS_L1_CC2),
       ("L2eq_rhs_CC2", _lib.RHS_L2_CC2)]


# Engine modes: the default lowering; every skinny contraction (one side with
# at most 32 rows) on the streaming VALU kernel; the same with every k-major
# one whose summed index runs a few contiguous elements on the short-run
# kernels (by default only from 512 elements per block on); no skinny path (everything
# small on the strided MFMA kernel); every batch-free contraction through the
# dense-GEMM path, outputs antisymmetric in two column letters computed for
# their ordered pairs at any summed length (by default only from K = 1024 on);
# the dense-GEMM path with its K dimension split over 3 partial products; the
# particle-particle ladders on all (i, j) rows instead of the i < j rows of
# the antisymmetric amplitudes; antisymmetric linear combinations summed at
# every element instead of the canonical ones; every dense-GEMM contraction
# with at most 512 A rows and 512 B rows (planes counted) on the dot-product
# kernel (by default only from K = 8192 on); every linear combination and
# outer product materialized (by default a linear combination used only by
# another is folded into it, and outer products of two matrices are formed
# where a linear combination adds them); linear combinations with product
# terms on the element-wise kernel (by default on the vectorized flat one when
# their plain terms are laid out like the output); every RK4 step on one
# stream (by default the T side and the Lambda side run on two); the GETT
# kernel for every skinny shape with each K offset read from its table (by
# default evenly stepped offsets are computed); linear combinations over a term
# contiguous along a leading axis on the element-wise kernel (by default that
# axis is exchanged with the second-to-last and the LDS-tiled kernel runs).
MODES = {"default": {}, "skinny": {"TDCC_SKINNY_MIN": "0"}, "gett": {"TDCC_SKINNY_MIN": "-1"},
         "skinny_runs": {"TDCC_SKINNY_MIN": "0", "TDCC_SKINNY_RUNS_MIN": "0"},
         "gemm": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_NPAIR_MIN_K": "0"},
         "gemm_splitk": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_GEMM_SPLITS": "3"},
         "full_rows": {"TDCC_IJ_PAIRS": "0"}, "plain_lincomb": {"TDCC_LINCOMB_ANTI": "0"},
         "dot": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_DOT_MIN_K": "0"},
         "no_inline": {"TDCC_INLINE_LINCOMB": "0", "TDCC_FOLD_OUTER": "0"},
         "lincomb_elementwise": {"TDCC_LINCOMB_FLAT_PROD": "0"}, "one_stream": {"TDCC_STREAMS": "1"},
         "gett_tables": {"TDCC_SKINNY_MIN": "-1", "TDCC_GETT_LINEAR_K": "0"},
         "lincomb_no_swap": {"TDCC_LINCOMB_SWAP": "0"}}
ENV_KEYS = ("TDCC_GEMM_MIN_FLOPS", "TDCC_GEMM_SPLITS", "TDCC_SKINNY_MIN", "TDCC_SKINNY_RUNS_MIN", "TDCC_IJ_PAIRS",
            "TDCC_NPAIR_MIN_K",
            "TDCC_LINCOMB_ANTI", "TDCC_DOT_MIN_K", "TDCC_INLINE_LINCOMB", "TDCC_FOLD_OUTER", "TDCC_LINCOMB_FLAT_PROD",
            "TDCC_STREAMS", "TDCC_GETT_LINEAR_K", "TDCC_LINCOMB_SWAP")


@pytest.fixture(scope="module", params=[(n, m) for m in MODES for n in ("o4v6", "c1_o10v4")],
                ids=lambda p: f"{p[0]}-{p[1]}")
def case(request):
    name, mode = request.param
    g = dict(np.load(os.path.join(HERE, "golden", f"golden_{name}.npz")))
    print(f"\n[case {name} engine={mode}]", flush=True)
    saved = {k: os.environ.get(k) for k in ENV_KEYS}
    for k in saved:
        os.environ.pop(k, None)
    os.environ.update(MODES[mode])
    ctx = _lib.Context(int(g["o"]), int(g["v"]))
    ctx.set_system(g["TEI"], g["F"], g["mu"])
    try:
        yield g, ctx
    finally:
        ctx.close()
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


@pytest.mark.parametrize("name", INTERMEDIATES)
def test_intermediate(case, name):
    g, ctx = case
    ref = g["out_" + name]
    got = ctx.intermediate(name, ref.shape, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"])
    assert np.abs(got - ref).max() <= TOL


@pytest.mark.parametrize("name,kind", RHS)
def test_rhs(case, name, kind):
    g, ctx = case
    got = ctx.rhs(kind, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"])
    ref = g["out_" + name]
    assert np.abs(got - ref).max() <= TOL


def test_dipole(case):
    g, ctx = case
    got = ctx.dipole(g["t1"], g["t2"], g["lam1"], g["lam2"])
    assert np.abs(got - g["out_dipole_moment"]).max() <= TOL


@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_rk4_trajectory(case, method):
    g, ctx = case
    ctx.set_state(g["t1"], g["t2"], g["lam1"], g["lam2"])
    n = int(g["rk_steps"])
    mu, ts = ctx.rk4_run(method, n, 0.0, float(g["rk_dt"]), float(g["rk_w0"]), float(g["rk_A"]))
    assert np.abs(mu - g[f"rk_{method}_mu"][1:]).max() <= TOL
    state = ctx.get_state(0)
    for k, key in enumerate(("t1", "t2", "lam1", "lam2")):
        assert np.abs(state[k] - g[f"rk_{method}_{key}"][-1]).max() <= TOL, key


def test_dot_kernel_row_tiles_match_gemm_path():
    """Dot-product contractions with more than 64 rows on a side run on several
    64-row tiles of the dot kernel.  At o = 8, v = 64 (up to 128 rows), with every
    batch-free contraction on the dense-GEMM path, the residuals computed with
    the dot-product shapes on the dot kernel must agree with the dense GEMMs to
    FP64 rounding."""
    from python_code_to_check_c_amd import synthetic
    o, v = 8, 64
    sysm = synthetic.SyntheticSystem(3, o, v, with_tei=False)
    rng = np.random.default_rng(11)

    def amp(shape):
        return 1e-2 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))

    t1, l1 = amp((o, v)), amp((o, v))
    t2, l2 = amp((o, o, v, v)), amp((o, o, v, v))
    for x in (t2, l2):
        x -= x.transpose(1, 0, 2, 3)
        x -= x.transpose(0, 1, 3, 2)
    Fc = sysm.F + 0j
    out = {}
    saved = {k: os.environ.get(k) for k in ("TDCC_DOT_MIN_K", "TDCC_GEMM_MIN_FLOPS")}
    try:
        os.environ["TDCC_GEMM_MIN_FLOPS"] = "0"
        for mode in ("0", "-1"):
            os.environ["TDCC_DOT_MIN_K"] = mode
            c = _lib.Context(o, v)
            try:
                c.set_system(None, np.asarray(sysm.F), np.asarray(sysm.mu))
                c.synth_tei(3, 0.05)
                out[mode] = [c.rhs(kind, Fc, t1, t2, l1, l2) for _, kind in RHS]
            finally:
                c.close()
    finally:
        for k, val in saved.items():
            if val is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = val
    for (name, _), x, y in zip(RHS, out["0"], out["-1"]):
        assert np.abs(x - y).max() <= 1e-12 * max(np.abs(y).max(), 1e-300), name


@pytest.mark.parametrize("kind,name", [(_lib.RHS_T2, "T2eq_rhs"), (_lib.RHS_L2, "lam2eq_rhs")])
def test_rhs_without_ij_antisymmetry(kind, name):
    """Amplitudes that are not antisymmetric in (i, j) take the full-row
    ladders; the result still matches the oracle's einsums on those inputs."""
    from oracle.cc_equations import CCEquations
    g = dict(np.load(os.path.join(HERE, "golden", "golden_o4v6.npz")))
    o = int(g["o"])
    rng = np.random.default_rng(7)
    t2 = g["t2"] + 0.05 * (rng.standard_normal(g["t2"].shape) + 1j * rng.standard_normal(g["t2"].shape))
    l2 = g["lam2"] + 0.05 * rng.standard_normal(g["lam2"].shape)
    assert np.abs(t2 + t2.transpose(1, 0, 2, 3)).max() > 1e-3
    eq = CCEquations(g["TEI"], o)
    ref = eq.T2eq_rhs(g["t1"], t2, g["Fc"]) if name == "T2eq_rhs" else \
        eq.lam2eq_rhs(g["t1"], t2, g["lam1"], l2, g["Fc"])
    ctx = _lib.Context(o, int(g["v"]))
    try:
        ctx.set_system(g["TEI"], g["F"], g["mu"])
        got = ctx.rhs(kind, g["Fc"], g["t1"], t2, g["lam1"], l2)
    finally:
        ctx.close()
    assert np.abs(got - ref).max() <= TOL
