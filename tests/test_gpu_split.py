"""Split-precision contractions (int8 digit planes, exact int32 accumulation) on the GPU.

Every batch-free contraction is forced through the dense-GEMM path
(TDCC_GEMM_MIN_FLOPS=0), so at these small shapes all rings, o^2 v^3 terms and
the ladders run on the digit GEMM.  The reference's outputs (golden vectors)
are the yardstick:
  * each operand element is truncated to 0.5 * 254^-S of its row scale
    (radix-254 digits), so residuals, the dipole and RK4 trajectories must
    agree with the reference to a relative error of at most 30 * 128^-S (a
    bar set for radix-128 digits, looser by (254/128)^S than radix 254 needs;
    the constant covers the sums of up to a few hundred terms at these shapes);
  * int32 accumulation is exact and the FP64 assembly has a fixed order, so a
    repeated evaluation is bitwise identical;
  * switching back to FP64 gives the FP64 results (1e-11).
"""
import os

import numpy as np
import pytest

from python_code_to_check_c_amd import _lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
RHS = [("T1eq_rhs", _lib.RHS_T1), ("T2eq_rhs", _lib.RHS_T2), ("lam_1eq_rhs", _lib.RHS_L1),
       ("lam2eq_rhs", _lib.RHS_L2), ("T2eq_rhs_CC2", _lib.RHS_T2_CC2), ("L1eq_rhs_CC2", _lib.RHS_L1_CC2),
       ("L2eq_rhs_CC2", _lib.RHS_L2_CC2)]


def bound(S):
    return 30.0 * 128.0 ** -S


# The digit split has single-pass kernels (rows held in registers) and
# three-pass kernels (partial maxima, scales, digits); "multipass" forces the
# latter everywhere.  "stream" keeps the default routing but sends every
# skinny contraction on an integral block through the split GEMM path (at
# large sizes only the blocks above 8M elements go there).  "default" also
# restricts outputs antisymmetric in two column letters to their ordered pairs
# at every summed length (at large sizes only from K = 1024 on).  "dot" sends
# the dot-product shapes (at most 64 A rows, 512 B rows) to the FP64 dot
# kernel at every summed length (at large sizes only from K = 8192 on).
# "chunked" cuts every dense-GEMM contraction into 3 K splits and gives each
# row-oriented operand with K > 256 one scale per row and split, split in one
# pass per chunk (at large sizes only for rows longer than 16384).  "m3"
# computes every complex x complex digit GEMM from three real products
# (re re, im im, (re + im)(re + im); at large sizes only the single-split
# ones above 512 x 512 complex outputs and 1e9 multiply-adds).  "tile64"
# runs every LDS-direct digit GEMM on 64-row tiles (at large sizes only those
# with at most 64 A rows).  "cols_resident" sends short-K column-oriented
# splits to the 64-row register-resident kernel whatever their row count (at
# large sizes only those above 2048 rows; below, one wave per row).
MODES = {"default": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_NPAIR_MIN_K": "0"},
         "multipass": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_SPLIT_MULTIPASS": "1"},
         "stream": {"TDCC_STATIC_STREAM_MIN": "0"},
         "dot": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_DOT_MIN_K": "0"},
         "chunked": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_GEMM_SPLITS": "3", "TDCC_SPLIT_CHUNKED": "1"},
         "m3": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_M3": "2"},
         "tile64": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_OZ_TILE64": "2"},
         "cols_resident": {"TDCC_GEMM_MIN_FLOPS": "0", "TDCC_WAVE_COLS_MAX_ROWS": "0"}}


@pytest.fixture(scope="module", params=[(n, m) for m in MODES for n in ("o4v6", "c1_o10v4")],
                ids=lambda p: f"{p[0]}-{p[1]}")
def case(request):
    name, mode = request.param
    g = dict(np.load(os.path.join(HERE, "golden", f"golden_{name}.npz")))
    print(f"\n[split case {name} {mode}]", flush=True)
    keys = ("TDCC_GEMM_MIN_FLOPS", "TDCC_SPLIT_MULTIPASS", "TDCC_STATIC_STREAM_MIN", "TDCC_NPAIR_MIN_K",
            "TDCC_DOT_MIN_K", "TDCC_GEMM_SPLITS", "TDCC_SPLIT_CHUNKED", "TDCC_M3", "TDCC_OZ_TILE64",
            "TDCC_WAVE_COLS_MAX_ROWS")
    saved = {k: os.environ.get(k) for k in keys}
    for k in keys:
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


@pytest.mark.parametrize("S", [2, 3, 4, 5, 6])
def test_residuals_within_truncation_bound(case, S):
    g, ctx = case
    ctx.set_precision(S)
    for name, kind in RHS:
        got = ctx.rhs(kind, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"])
        ref = g["out_" + name]
        rel = np.abs(got - ref).max() / np.abs(ref).max()
        assert rel <= bound(S), (name, rel)
    got = ctx.dipole(g["t1"], g["t2"], g["lam1"], g["lam2"])
    ref = g["out_dipole_moment"]
    assert np.abs(got - ref).max() <= bound(S) * np.abs(ref).max()


@pytest.mark.parametrize("S", [3, 4, 6])
@pytest.mark.parametrize("method", ["CCSD", "CC2"])
def test_rk4_trajectory_within_bound(case, S, method):
    g, ctx = case
    ctx.set_precision(S)
    ctx.set_state(g["t1"], g["t2"], g["lam1"], g["lam2"])
    mu, _ = ctx.rk4_run(method, int(g["rk_steps"]), 0.0, float(g["rk_dt"]), float(g["rk_w0"]), float(g["rk_A"]))
    ref = g[f"rk_{method}_mu"][1:]
    assert np.abs(mu - ref).max() <= bound(S) * np.abs(ref).max()


def test_split_is_deterministic(case):
    g, ctx = case
    ctx.set_precision(4)
    a = [ctx.rhs(kind, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"]) for _, kind in RHS]
    b = [ctx.rhs(kind, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"]) for _, kind in RHS]
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_shared_digits_are_bitwise_neutral(case):
    """Launches that split one computed operand in one layout share its digit
    planes (cut once, at the first of them); without sharing every launch cuts
    its own copy of the same digits, so the results must be bitwise equal."""
    g, ctx = case
    ctx.set_precision(4)

    def evaluate(c):
        out = [c.rhs(kind, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"]) for _, kind in RHS]
        c.set_state(g["t1"], g["t2"], g["lam1"], g["lam2"])
        mu, _ = c.rk4_run("CCSD", int(g["rk_steps"]), 0.0, float(g["rk_dt"]), float(g["rk_w0"]), float(g["rk_A"]))
        return out + [mu]

    shared = evaluate(ctx)
    os.environ["TDCC_SHARE_DIGITS"] = "0"
    try:
        c2 = _lib.Context(int(g["o"]), int(g["v"]))
        try:
            c2.set_system(g["TEI"], g["F"], g["mu"])
            c2.set_precision(4)
            own = evaluate(c2)
        finally:
            c2.close()
    finally:
        os.environ.pop("TDCC_SHARE_DIGITS", None)
    for x, y in zip(shared, own):
        assert np.array_equal(x, y)


def test_persistent_digit_gemm_is_bitwise_neutral():
    """The 4-digit GEMM runs persistent blocks that loop over the tiles
    (TDCC_OZ_PERSIST blocks; 0: one block per tile).  A tile's arithmetic does
    not depend on the block that computes it, so at a size with 64 tiles per
    GEMM (o = 8, v = 64, every contraction on the GEMM path) 8 looping blocks
    and one block per tile must give bitwise equal residuals."""
    from python_code_to_check_c_amd import synthetic
    o, v = 8, 64
    sysm = synthetic.SyntheticSystem(3, o, v, with_tei=False)
    rng = np.random.default_rng(5)

    def amp(shape):
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        return 1e-2 * x

    t1, l1 = amp((o, v)), amp((o, v))
    t2, l2 = amp((o, o, v, v)), amp((o, o, v, v))
    for x in (t2, l2):
        x -= x.transpose(1, 0, 2, 3)
        x -= x.transpose(0, 1, 3, 2)
    Fc = sysm.F + 0j
    out = {}
    saved = {k: os.environ.get(k) for k in ("TDCC_OZ_PERSIST", "TDCC_GEMM_MIN_FLOPS")}
    try:
        os.environ["TDCC_GEMM_MIN_FLOPS"] = "0"
        for blocks in ("8", "0"):
            os.environ["TDCC_OZ_PERSIST"] = blocks
            c = _lib.Context(o, v)
            try:
                c.set_system(None, np.asarray(sysm.F), np.asarray(sysm.mu))
                c.synth_tei(3, 0.05)
                c.set_precision(4)
                out[blocks] = [c.rhs(kind, Fc, t1, t2, l1, l2) for _, kind in RHS]
            finally:
                c.close()
    finally:
        for k, val in saved.items():
            if val is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = val
    for (name, _), x, y in zip(RHS, out["8"], out["0"]):
        assert np.array_equal(x, y), name


@pytest.mark.parametrize("digits", [4, 5])
def test_small_tile_digit_gemm_is_bitwise_neutral(digits):
    """The LDS-direct digit GEMM takes 64-row tiles for A operands of at most
    64 rows (TDCC_OZ_TILE64; 2: every launch, 0: always 128 rows).  The
    integer sums and their FP64 assembly do not depend on the tile shape, so
    at o = 8, v = 64 (every contraction on the GEMM path) both tilings give
    bitwise equal residuals, for 4 and 5 digits, and with three real products
    for the complex ones."""
    from python_code_to_check_c_amd import synthetic
    o, v = 8, 64
    sysm = synthetic.SyntheticSystem(3, o, v, with_tei=False)
    rng = np.random.default_rng(6)

    def amp(shape):
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        return 1e-2 * x

    t1, l1 = amp((o, v)), amp((o, v))
    t2, l2 = amp((o, o, v, v)), amp((o, o, v, v))
    for x in (t2, l2):
        x -= x.transpose(1, 0, 2, 3)
        x -= x.transpose(0, 1, 3, 2)
    Fc = sysm.F + 0j
    out = {}
    keys = ("TDCC_OZ_TILE64", "TDCC_GEMM_MIN_FLOPS", "TDCC_M3")
    saved = {k: os.environ.get(k) for k in keys}
    try:
        os.environ["TDCC_GEMM_MIN_FLOPS"] = "0"
        os.environ["TDCC_M3"] = "2"
        for mode in ("2", "0"):
            os.environ["TDCC_OZ_TILE64"] = mode
            c = _lib.Context(o, v)
            try:
                c.set_system(None, np.asarray(sysm.F), np.asarray(sysm.mu))
                c.synth_tei(3, 0.05)
                c.set_precision(digits)
                out[mode] = [c.rhs(kind, Fc, t1, t2, l1, l2) for _, kind in RHS]
            finally:
                c.close()
    finally:
        for k, val in saved.items():
            if val is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = val
    for (name, _), x, y in zip(RHS, out["2"], out["0"]):
        assert np.array_equal(x, y), name


def test_back_to_fp64(case):
    g, ctx = case
    ctx.set_precision(5)
    ctx.rhs(_lib.RHS_T2, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"])
    ctx.set_precision(0)
    for name, kind in RHS:
        got = ctx.rhs(kind, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"])
        assert np.abs(got - g["out_" + name]).max() <= 1e-11, name


def test_rejects_bad_digit_count(case):
    _, ctx = case
    for bad in (1, 7, -1):
        with pytest.raises(_lib.TdccError):
            ctx.set_precision(bad)


@pytest.mark.parametrize("digits", [4, 5])
def test_fused_three_plane_split_is_bitwise_neutral(digits):
    """Three-plane operands of the 3-product complex GEMMs (re, im, re + im)
    are split from one read of each base row (TDCC_SPLIT3_FUSED=0: plane by
    plane).  The digits and scales are the same numbers, so at o = 8, v = 64
    with every complex x complex GEMM on three products (summed length 512:
    the row-resident and the three-pass column splits) the residuals are
    bitwise equal."""
    from python_code_to_check_c_amd import synthetic
    o, v = 8, 64
    sysm = synthetic.SyntheticSystem(3, o, v, with_tei=False)
    rng = np.random.default_rng(8)

    def amp(shape):
        return 1e-2 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))

    t1, l1 = amp((o, v)), amp((o, v))
    t2, l2 = amp((o, o, v, v)), amp((o, o, v, v))
    for x in (t2, l2):
        x -= x.transpose(1, 0, 2, 3)
        x -= x.transpose(0, 1, 3, 2)
    Fc = sysm.F + 0j
    out = {}
    keys = ("TDCC_SPLIT3_FUSED", "TDCC_GEMM_MIN_FLOPS", "TDCC_M3")
    saved = {k: os.environ.get(k) for k in keys}
    try:
        os.environ["TDCC_GEMM_MIN_FLOPS"] = "0"
        os.environ["TDCC_M3"] = "2"
        for mode in ("1", "0"):
            os.environ["TDCC_SPLIT3_FUSED"] = mode
            c = _lib.Context(o, v)
            try:
                c.set_system(None, np.asarray(sysm.F), np.asarray(sysm.mu))
                c.synth_tei(3, 0.05)
                c.set_precision(digits)
                out[mode] = [c.rhs(kind, Fc, t1, t2, l1, l2) for _, kind in RHS]
            finally:
                c.close()
    finally:
        for k, val in saved.items():
            if val is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = val
    for (name, _), x, y in zip(RHS, out["1"], out["0"]):
        assert np.array_equal(x, y), name


@pytest.mark.parametrize("digits", [0, 5])
def test_lincomb_axis_swap_is_bitwise_neutral(digits):
    """A linear combination whose term is contiguous along a leading axis (a
    permuted integral block) runs on the LDS-tiled kernel with that axis
    exchanged for the second-to-last one (TDCC_LINCOMB_SWAP=0: element-wise
    kernel).  Each element is the same sum in the same term order, so at
    o = 8, v = 64 the residuals and one profiled step are bitwise equal, and
    the swap moves at least one launch onto the tiled kernel."""
    from python_code_to_check_c_amd import synthetic
    o, v = 8, 64
    sysm = synthetic.SyntheticSystem(3, o, v, with_tei=False)
    rng = np.random.default_rng(12)

    def amp(shape):
        return 1e-2 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))

    t1, l1 = amp((o, v)), amp((o, v))
    t2, l2 = amp((o, o, v, v)), amp((o, o, v, v))
    for x in (t2, l2):
        x -= x.transpose(1, 0, 2, 3)
        x -= x.transpose(0, 1, 3, 2)
    Fc = sysm.F + 0j
    out, state, tiled = {}, {}, {}
    saved = os.environ.get("TDCC_LINCOMB_SWAP")
    try:
        for mode in ("1", "0"):
            os.environ["TDCC_LINCOMB_SWAP"] = mode
            c = _lib.Context(o, v)
            try:
                c.set_system(None, np.asarray(sysm.F), np.asarray(sysm.mu))
                c.synth_tei(3, 0.05)
                c.set_precision(digits)
                out[mode] = [c.rhs(kind, Fc, t1, t2, l1, l2) for _, kind in RHS]
                c.set_state(t1, t2, l1, l2)
                names = [r[0] for r in c.profile_step("CCSD", 0.0, 1e-3, 0.923, 2e-3)]
                tiled[mode] = sum(n == "lincomb_tiled_kernel" for n in names)
                state[mode] = c.get_state(0)
            finally:
                c.close()
    finally:
        if saved is None:
            os.environ.pop("TDCC_LINCOMB_SWAP", None)
        else:
            os.environ["TDCC_LINCOMB_SWAP"] = saved
    for (name, _), x, y in zip(RHS, out["1"], out["0"]):
        assert np.array_equal(x, y), name
    for x, y in zip(state["1"], state["0"]):
        assert np.array_equal(x, y)
    assert tiled["1"] > tiled["0"], tiled


@pytest.mark.parametrize("digits", [0, 5])
@pytest.mark.parametrize("batch", ["single", "stacked", "side_by_side"])
def test_two_stream_step_is_bitwise_neutral(digits, batch):
    """An RK4 step runs its T side and its Lambda side (the Lambda chains read
    the step-start t) on two streams joined before the update, with the
    buffers the Lambda side reads from the T side held through the step;
    TDCC_STREAMS=1 keeps one stream.  The same kernels compute the same
    values in either order, so three steps of CCSD give bitwise equal
    amplitudes and dipoles, in FP64 and with 5 digits, for one trajectory and
    for a batch of two (stacked and side by side)."""
    from python_code_to_check_c_amd import synthetic
    o, v = 8, 64
    sysm = synthetic.SyntheticSystem(3, o, v, with_tei=False)
    rng = np.random.default_rng(11)

    def amp(shape):
        return 1e-2 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))

    states = []
    for _ in range(2):
        t1, l1 = amp((o, v)), amp((o, v))
        t2, l2 = amp((o, o, v, v)), amp((o, o, v, v))
        for x in (t2, l2):
            x -= x.transpose(1, 0, 2, 3)
            x -= x.transpose(0, 1, 3, 2)
        states.append((t1, t2, l1, l2))
    out = {}
    keys = ("TDCC_STREAMS", "TDCC_GEMM_MIN_FLOPS", "TDCC_STACK")
    saved = {k: os.environ.get(k) for k in keys}
    try:
        os.environ["TDCC_GEMM_MIN_FLOPS"] = "0"
        os.environ["TDCC_STACK"] = "0" if batch == "side_by_side" else "1"
        for mode in ("2", "1"):
            os.environ["TDCC_STREAMS"] = mode
            c = _lib.Context(o, v)
            try:
                c.set_system(None, np.asarray(sysm.F), np.asarray(sysm.mu))
                c.synth_tei(3, 0.05)
                c.set_precision(digits)
                if batch == "single":
                    c.set_state(*states[0])
                    mu, _ = c.rk4_run("CCSD", 3, 0.0, 0.05, 0.3, 0.02)
                    out[mode] = [mu, *c.get_state()]
                else:
                    c.set_batch(2)
                    for b in range(2):
                        c.set_state_slot(b, *states[b])
                    mu, _ = c.rk4_run_batch("CCSD", 3, 0.0, 0.05, [2, 0], [0.3, 0.25], [0.02, 0.03])
                    out[mode] = [mu] + [x for b in range(2) for x in c.get_state_slot(b)]
            finally:
                c.close()
    finally:
        for k, val in saved.items():
            if val is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = val
    assert np.abs(out["2"][0]).max() > 0
    for x, y in zip(out["2"], out["1"]):
        assert np.array_equal(x, y)
