"""Device-memory entry points (tdcc_set_state_device, tdcc_get_state_device,
tdcc_rhs_device) against their host counterparts.

The device variants read and write GPU memory and never copy through the
host, but run the same programs: every residual of the golden cases must
equal tdcc_rhs bit for bit and match the reference's golden vectors to 1e-11;
a state set from device memory must read back bit for bit and propagate
exactly like one set from host arrays; the GPU antisymmetry test must pick
the same (i, j) shortcut as the host test (an amplitude that is not
antisymmetric takes the full-row programs on both); host memory handed to a
device entry point is refused with an error.  In this process the buffers
are DeviceArrays (the library's own HIP runtime); GPU torch tensors are
checked in a fresh process that initializes torch's GPU before the library
loads, so that both share torch's runtime.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

from python_code_to_check_c_amd import _lib

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TOL = 1e-11
RHS = [("T1eq_rhs", _lib.RHS_T1), ("T2eq_rhs", _lib.RHS_T2), ("lam_1eq_rhs", _lib.RHS_L1),
       ("lam2eq_rhs", _lib.RHS_L2), ("T2eq_rhs_CC2", _lib.RHS_T2_CC2), ("L1eq_rhs_CC2", _lib.RHS_L1_CC2),
       ("L2eq_rhs_CC2", _lib.RHS_L2_CC2)]
KEYS = ("t1", "t2", "lam1", "lam2")


@pytest.fixture(scope="module", params=["o4v6", "c1_o10v4"])
def case(request):
    g = dict(np.load(os.path.join(HERE, "golden", f"golden_{request.param}.npz")))
    print(f"\n[device io case {request.param}]", flush=True)
    ctx = _lib.Context(int(g["o"]), int(g["v"]))
    ctx.set_system(g["TEI"], g["F"], g["mu"])
    try:
        yield g, ctx
    finally:
        ctx.close()


def _dev(x):
    return _lib.DeviceArray.from_numpy(x)


@pytest.mark.parametrize("name,kind", RHS)
def test_rhs_device_matches_host(case, name, kind):
    g, ctx = case
    host = ctx.rhs(kind, g["Fc"], g["t1"], g["t2"], g["lam1"], g["lam2"])
    got = ctx.rhs_device(kind, _dev(g["Fc"]), *[_dev(g[k]) for k in KEYS]).to_numpy()
    assert np.array_equal(got, host)
    assert np.abs(got - g["out_" + name]).max() <= TOL


def test_rhs_device_without_ij_antisymmetry(case):
    """A t2 that breaks (i, j) antisymmetry: both tests see it, both take the
    full-row program, and the results agree bit for bit."""
    g, ctx = case
    rng = np.random.default_rng(3)
    t2 = g["t2"] + 1e-3 * (rng.standard_normal(g["t2"].shape) + 1j * rng.standard_normal(g["t2"].shape))
    for kind in (_lib.RHS_T2, _lib.RHS_L2):
        host = ctx.rhs(kind, g["Fc"], g["t1"], t2, g["lam1"], g["lam2"])
        dev = ctx.rhs_device(kind, _dev(g["Fc"]), _dev(g["t1"]), _dev(t2), _dev(g["lam1"]), _dev(g["lam2"]))
        assert np.array_equal(dev.to_numpy(), host)


def test_state_device_round_trip_and_step(case):
    g, ctx = case
    n = int(g["rk_steps"])
    args = (0.0, float(g["rk_dt"]), float(g["rk_w0"]), float(g["rk_A"]))
    ctx.set_state(*[g[k] for k in KEYS])
    mu_host, _ = ctx.rk4_run("CCSD", n, *args)
    end_host, base_host = ctx.get_state(0), ctx.get_state(1)
    ctx.set_state_device(*[_dev(g[k]) for k in KEYS])
    for x, k in zip(ctx.get_state(0), KEYS):
        assert np.array_equal(x, g[k])
    for x, k in zip(ctx.get_state_device(0), KEYS):
        assert np.array_equal(x.to_numpy(), g[k])
    mu_dev, _ = ctx.rk4_run("CCSD", n, *args)
    assert np.array_equal(mu_dev, mu_host)
    assert np.abs(mu_dev - g["rk_CCSD_mu"][1:]).max() <= TOL
    for a, b in zip(ctx.get_state_device(0), end_host):
        assert np.array_equal(a.to_numpy(), b)
    for a, b in zip(ctx.get_state_device(1), base_host):
        assert np.array_equal(a.to_numpy(), b)


def test_device_entry_refuses_host_memory(case):
    g, ctx = case
    host = [np.ascontiguousarray(g[k], dtype=np.complex128) for k in KEYS]
    rc = ctx.lib.tdcc_set_state_device(ctx._h, *[ctypes.c_void_p(x.ctypes.data) for x in host])
    assert rc != 0
    assert b"not device memory" in ctx.lib.tdcc_last_error(ctx._h)
    with pytest.raises(TypeError):
        ctx.set_state_device(*host)


TORCH_SCRIPT = r"""
import sys
import numpy as np
import torch
torch.zeros(1, device="cuda:0")
torch.cuda.synchronize()
sys.path.insert(0, sys.argv[1])
from python_code_to_check_c_amd import _lib
g = dict(np.load(sys.argv[2]))
ctx = _lib.Context(int(g["o"]), int(g["v"]))
ctx.set_system(g["TEI"], g["F"], g["mu"])
gpu = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).to("cuda:0")
keys = ("t1", "t2", "lam1", "lam2")
names = ("T1eq_rhs", "T2eq_rhs", "lam_1eq_rhs", "lam2eq_rhs", "T2eq_rhs_CC2", "L1eq_rhs_CC2", "L2eq_rhs_CC2")
bad = []
for kind, name in enumerate(names):
    host = ctx.rhs(kind, g["Fc"], *[g[k] for k in keys])
    dev = ctx.rhs_device(kind, gpu(g["Fc"]), *[gpu(g[k]) for k in keys])
    if not isinstance(dev, torch.Tensor) or not np.array_equal(dev.cpu().numpy(), host):
        bad.append(name)
ctx.set_state_device(*[gpu(g[k]) for k in keys])
back = ctx.get_state_device(0, like=gpu(g["t1"]))
if not all(np.array_equal(a.cpu().numpy(), g[k]) for a, k in zip(back, keys)):
    bad.append("state")
print("MISMATCH " + " ".join(bad) if bad else "TORCH_OK", flush=True)
"""


def test_torch_tensors_on_shared_runtime():
    """GPU torch tensors in and out, in a process where torch's GPU comes up
    first and the library shares its HIP runtime: bit for bit the host
    entry points' results."""
    golden = os.path.join(HERE, "golden", "golden_o4v6.npz")
    r = subprocess.run([sys.executable, "-c", TORCH_SCRIPT, ROOT, golden], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "TORCH_OK" in r.stdout, r.stdout[-2000:]
