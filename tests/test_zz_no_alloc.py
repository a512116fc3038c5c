"""tdcc_dipole_device_noalloc (no device allocation; scratch in a static
device array) against tdcc_dipole on the same amplitudes: two calls in a row
on one GPU, the second for a smaller system, each bit for bit equal to the
allocating entry point's result.  Device buffers for the inputs come from the
HIP runtime the library itself uses (hipMalloc through the library's handle)."""
import ctypes

import numpy as np
import pytest

from python_code_to_check_c_amd import _lib, synthetic

pytestmark = pytest.mark.gpu
H2D, D2H = 1, 2


class DeviceBuffer:
    """Device memory from the library's own HIP runtime (test plumbing)."""

    def __init__(self, host):
        lib = _lib.load()
        self.lib = lib
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        self.nbytes = host.nbytes
        self.ptr = ctypes.c_void_p()
        assert lib.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        self.upload(host)

    def upload(self, host):
        a = np.ascontiguousarray(host)
        assert a.nbytes == self.nbytes
        assert self.lib.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D) == 0

    def download(self, dtype):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr.value:
            self.lib.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _system(o, v, cid):
    s = synthetic.SyntheticSystem(cid, o, v, tei_sigma=0.02)
    rng = np.random.default_rng(100 + o + v)

    def amp(shape):
        x = 0.05 * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        if len(shape) == 4:
            x = x - x.transpose(1, 0, 2, 3)
            x = x - x.transpose(0, 1, 3, 2)
        return x

    return s, (amp((o, v)), amp((o, o, v, v)), amp((o, v)), amp((o, o, v, v)))


def test_noalloc_dipole_matches_allocating_entry_point():
    ctxs, bufs = [], []
    try:
        for (o, v, cid) in ((10, 38, 2), (4, 6, 11)):  # second call: a smaller system
            print(f"\n[no-alloc dipole o={o} v={v}]", flush=True)
            s, amps = _system(o, v, cid)
            ctx = _lib.Context(o, v)
            ctxs.append(ctx)
            ctx.set_system(s.TEI, s.F, s.mu)
            ref = ctx.dipole(*amps)
            dev = [DeviceBuffer(a) for a in amps]
            out = DeviceBuffer(np.full(6, np.nan))
            bufs += dev + [out]
            ctx.dipole_device_noalloc(*[d.addr for d in dev], out.addr)
            got = out.download(np.float64).view(np.complex128)
            assert np.array_equal(got, ref), (o, v, got, ref)
            # again after other work of the context: the same bits
            ctx.dipole(*amps)
            out.upload(np.zeros(6))
            ctx.dipole_device_noalloc(*[d.addr for d in dev], out.addr)
            assert np.array_equal(out.download(np.float64).view(np.complex128), ref)
    finally:
        for b in bufs:
            b.free()
        for c in ctxs:
            c.close()


def test_noalloc_dipole_rejects_bad_arguments():
    ctx = _lib.Context(4, 6)
    try:
        s, amps = _system(4, 6, 11)
        ctx.set_system(s.TEI, s.F, s.mu)
        rc = ctx.lib.tdcc_dipole_device_noalloc(ctx._h, None, None, None, None, None)
        assert rc != 0 and b"required" in ctx.lib.tdcc_last_error(ctx._h)
    finally:
        ctx.close()
