"""ctypes binding of libtdcc_hip.so (the C ABI declared in include/tdcc.h).

There is no CPU fallback: if the library is missing or no HIP device is
available, the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtdcc_hip.so")

# Every symbol include/tdcc.h declares.
EXPORTS = (
    "tdcc_create", "tdcc_destroy", "tdcc_last_error", "tdcc_set_system", "tdcc_synth_tei",
    "tdcc_set_precision", "tdcc_set_state", "tdcc_get_state", "tdcc_rhs", "tdcc_intermediate", "tdcc_dipole",
    "tdcc_rk4_run", "tdcc_profile_step", "tdcc_checkpoint", "tdcc_step_info",
    "tdcc_set_batch", "tdcc_set_state_slot", "tdcc_get_state_slot", "tdcc_rk4_run_batch",
    "tdcc_dipole_gradient", "tdcc_write_restart_block", "tdcc_io_error",
    "tdcc_dipole_device_noalloc", "tdcc_read_restart_block", "tdcc_write_state_file",
    "tdcc_read_state_file_header", "tdcc_read_state_file",
    "tdcc_set_state_device", "tdcc_get_state_device", "tdcc_rhs_device",
)

RHS_T1, RHS_T2, RHS_L1, RHS_L2, RHS_T2_CC2, RHS_L1_CC2, RHS_L2_CC2 = range(7)
METHODS = {"CCSD": 0, "CC2": 1}


class TdccError(RuntimeError):
    pass


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TdccError(f"{LIB_PATH} is missing; build it with "
                        "`python -m python_code_to_check_c_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    i64p = ctypes.POINTER(ctypes.c_int64)
    sig = {
        "tdcc_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]),
        "tdcc_destroy": (None, [vp]),
        "tdcc_last_error": (ctypes.c_char_p, [vp]),
        "tdcc_set_system": (ctypes.c_int, [vp, _dp, _dp, _dp]),
        "tdcc_synth_tei": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_double]),
        "tdcc_set_precision": (ctypes.c_int, [vp, ctypes.c_int]),
        "tdcc_set_state": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp]),
        "tdcc_get_state": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
        "tdcc_rhs": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "tdcc_intermediate": (ctypes.c_int, [vp, ctypes.c_char_p, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int64]),
        "tdcc_dipole": (ctypes.c_int, [vp, _dp, _dp, _dp, _dp, _dp]),
        "tdcc_dipole_device_noalloc": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
        "tdcc_rk4_run": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp]),
        "tdcc_profile_step": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                             ctypes.c_int64]),
        "tdcc_checkpoint": (ctypes.c_int, [vp, ctypes.c_int]),
        "tdcc_step_info": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, i64p, i64p]),
        "tdcc_dipole_gradient": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp,
                                                ctypes.c_int, _dp, _dp, _dp, _dp]),
        "tdcc_set_batch": (ctypes.c_int, [vp, ctypes.c_int]),
        "tdcc_set_state_slot": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
        "tdcc_get_state_slot": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
        "tdcc_rk4_run_batch": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.POINTER(ctypes.c_int), _dp, _dp, ctypes.c_int, _dp, _dp]),
        "tdcc_write_restart_block": (ctypes.c_int, [_dp, ctypes.c_int, i64p, ctypes.c_int, ctypes.c_char_p]),
        "tdcc_io_error": (ctypes.c_char_p, []),
        "tdcc_read_restart_block": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, i64p, _dp]),
        "tdcc_write_state_file": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]),
        "tdcc_read_state_file_header": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int),
                                                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                                       ctypes.POINTER(ctypes.c_int), _dp]),
        "tdcc_read_state_file": (ctypes.c_int, [ctypes.c_char_p, _dp, _dp, _dp, _dp, _dp]),
        "tdcc_set_state_device": (ctypes.c_int, [vp, vp, vp, vp, vp]),
        "tdcc_get_state_device": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, vp, vp]),
        "tdcc_rhs_device": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _require_one_hip_runtime():
    """GPU torch tensors are usable only when the library runs on torch's HIP
    runtime: torch (with its GPU initialized) must be loaded before the first
    Context, so that the library binds to torch's runtime instead of loading
    the system one beside it."""
    with open("/proc/self/maps") as f:
        runtimes = {ln.split()[-1] for ln in f if "libamdhip64" in ln}
    if len(runtimes) != 1:
        raise TdccError("GPU torch tensors need the library on torch's HIP runtime; two are loaded "
                        f"({sorted(runtimes)}): import torch and initialize its GPU before the first "
                        "Context, or pass DeviceArray buffers")


class DeviceArray:
    """Device memory from the library's own HIP runtime holding an array
    (complex128 amplitudes and F for the *_device entry points), for callers
    without torch."""

    def __init__(self, shape, dtype=np.complex128):
        lib = load()
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        self._lib = lib
        self.shape = tuple(int(n) for n in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = ctypes.c_void_p()
        if lib.hipMalloc(ctypes.byref(self._p), max(self.nbytes, 1)) != 0:
            raise TdccError(f"hipMalloc of {self.nbytes} bytes failed")

    @classmethod
    def from_numpy(cls, a, dtype=np.complex128):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, dtype)
        if d._lib.hipMemcpy(d._p, a.ctypes.data, d.nbytes, 1) != 0:  # host to device
            raise TdccError("hipMemcpy to the device failed")
        return d

    def to_numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self._lib.hipMemcpy(out.ctypes.data, self._p, self.nbytes, 2) != 0:  # device to host
            raise TdccError("hipMemcpy from the device failed")
        return out

    @property
    def ptr(self):
        return self._p.value

    def free(self):
        if self._p.value:
            self._lib.hipFree(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _c(x):
    a = np.ascontiguousarray(x, dtype=np.complex128)
    return a, a.ctypes.data_as(_dp)


def _r(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _opt(x):
    if x is None:
        return None, None
    return _c(x)


class Context:
    """One device context: integrals, Fock, dipoles and the propagated state."""

    def __init__(self, nocc, nvir, device=0):
        self.lib = load()
        self.device = int(device)
        self.o, self.v, self.n = int(nocc), int(nvir), int(nocc) + int(nvir)
        h = ctypes.c_void_p()
        if self.lib.tdcc_create(self.o, self.v, int(device), ctypes.byref(h)) != 0 or not h.value:
            raise TdccError(f"tdcc_create({nocc}, {nvir}, device={device}) failed "
                            "(no HIP device, or out of device memory)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.tdcc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.tdcc_last_error(self._h)
            raise TdccError(f"{what}: {msg.decode() if msg else 'error'}")

    # ------------------------------------------------------------------ setup
    def set_system(self, TEI, F, mu):
        n = self.n
        F_, pF = _r(F)
        mu_, pmu = _r(mu)
        if F_.shape != (n, n) or mu_.shape != (3, n, n):
            raise ValueError("F must be (n,n) and mu (3,n,n)")
        if TEI is None:
            self._check(self.lib.tdcc_set_system(self._h, None, pF, pmu), "tdcc_set_system")
            return
        T_, pT = _r(TEI)
        if T_.shape != (n, n, n, n):
            raise ValueError("TEI must be (n,n,n,n)")
        self._check(self.lib.tdcc_set_system(self._h, pT, pF, pmu), "tdcc_set_system")

    def synth_tei(self, config_id, sigma):
        self._check(self.lib.tdcc_synth_tei(self._h, int(config_id), float(sigma)), "tdcc_synth_tei")

    def set_precision(self, digits):
        """0: FP64 contractions; 2..6: split precision with that many int8
        digits per operand for the large contractions (see include/tdcc.h)."""
        self._check(self.lib.tdcc_set_precision(self._h, int(digits)), "tdcc_set_precision")
        self.digits = int(digits)

    def _shape_ok(self, t1, t2):
        if t1.shape != (self.o, self.v) or t2.shape != (self.o, self.o, self.v, self.v):
            raise ValueError(f"amplitude shapes {t1.shape}, {t2.shape} do not match o={self.o}, v={self.v}")

    def set_state(self, t1, t2, lam1, lam2):
        a, pa = _c(t1)
        b, pb = _c(t2)
        c, pc = _c(lam1)
        d, pd = _c(lam2)
        self._shape_ok(a, b)
        self._shape_ok(c, d)
        self._check(self.lib.tdcc_set_state(self._h, pa, pb, pc, pd), "tdcc_set_state")

    # ------------------------------------------------- device-resident arrays
    def _dev(self, x, shape, what):
        """Device address of `x`: a DeviceArray, or a contiguous complex128 GPU
        torch tensor on the context's GPU (torch's stream synchronized first),
        of the given shape."""
        if isinstance(x, DeviceArray):
            if x.shape != tuple(shape) or x.dtype != np.complex128:
                raise ValueError(f"{what} is a {x.dtype} DeviceArray of shape {x.shape}, "
                                 f"expected complex128 {tuple(shape)}")
            return ctypes.c_void_p(x.ptr)
        import torch
        if not isinstance(x, torch.Tensor) or not x.is_cuda:
            raise TypeError(f"{what} must be a DeviceArray or a GPU torch tensor")
        if x.dtype != torch.complex128:
            raise TypeError(f"{what} must be complex128, got {x.dtype}")
        if tuple(x.shape) != tuple(shape):
            raise ValueError(f"{what} has shape {tuple(x.shape)}, expected {tuple(shape)}")
        if not x.is_contiguous():
            raise ValueError(f"{what} must be contiguous")
        if x.device.index != self.device:
            raise ValueError(f"{what} is on cuda:{x.device.index}, the context on cuda:{self.device}")
        _require_one_hip_runtime()
        torch.cuda.current_stream(x.device).synchronize()
        return ctypes.c_void_p(x.data_ptr())

    def _empty_like(self, ref, shape):
        if isinstance(ref, DeviceArray):
            return DeviceArray(shape), None
        import torch
        t = torch.empty(shape, dtype=torch.complex128, device=ref.device)
        return t, ctypes.c_void_p(t.data_ptr())

    def _shapes(self):
        o, v = self.o, self.v
        return (o, v), (o, o, v, v)

    def set_state_device(self, t1, t2, lam1, lam2):
        """set_state from device memory (DeviceArray or GPU torch tensors,
        complex128): no host copy."""
        s1, s2 = self._shapes()
        ptrs = [self._dev(x, sh, nm) for x, sh, nm in
                ((t1, s1, "t1"), (t2, s2, "t2"), (lam1, s1, "lam1"), (lam2, s2, "lam2"))]
        self._check(self.lib.tdcc_set_state_device(self._h, *ptrs), "tdcc_set_state_device")

    def get_state_device(self, which=0, like=None):
        """get_state into new device arrays: GPU torch tensors on `like`'s
        device when `like` is a torch tensor, else DeviceArrays."""
        s1, s2 = self._shapes()
        shapes = (s1, s2, s1, s2)
        if like is not None and not isinstance(like, DeviceArray):
            import torch
            _require_one_hip_runtime()
            out = [torch.empty(sh, dtype=torch.complex128, device=like.device) for sh in shapes]
            ptrs = [ctypes.c_void_p(x.data_ptr()) for x in out]
            torch.cuda.current_stream(like.device).synchronize()
        else:
            out = [DeviceArray(sh) for sh in shapes]
            ptrs = [ctypes.c_void_p(x.ptr) for x in out]
        self._check(self.lib.tdcc_get_state_device(self._h, int(which), *ptrs), "tdcc_get_state_device")
        return tuple(out)

    def rhs_device(self, kind, F, t1, t2, lam1=None, lam2=None):
        """rhs on device memory (DeviceArray or GPU torch tensors, complex128,
        result of t1's kind): no host copies."""
        s1, s2 = self._shapes()
        lam = kind in (RHS_L1, RHS_L2, RHS_L1_CC2, RHS_L2_CC2)
        if lam and (lam1 is None or lam2 is None):
            raise ValueError("lam1 and lam2 are required for a Lambda right-hand side")
        pF = self._dev(F, (self.n, self.n), "F")
        p1, p2 = self._dev(t1, s1, "t1"), self._dev(t2, s2, "t2")
        pl1 = self._dev(lam1, s1, "lam1") if lam else None
        pl2 = self._dev(lam2, s2, "lam2") if lam else None
        out, po = self._empty_like(t1, s1 if kind in (RHS_T1, RHS_L1, RHS_L1_CC2) else s2)
        if po is None:
            po = ctypes.c_void_p(out.ptr)
        self._check(self.lib.tdcc_rhs_device(self._h, int(kind), pF, p1, p2, pl1, pl2, po), "tdcc_rhs_device")
        return out

    def get_state(self, which=0):
        o, v = self.o, self.v
        out = [np.empty((o, v), complex), np.empty((o, o, v, v), complex),
               np.empty((o, v), complex), np.empty((o, o, v, v), complex)]
        ptrs = [x.ctypes.data_as(_dp) for x in out]
        self._check(self.lib.tdcc_get_state(self._h, int(which), *ptrs), "tdcc_get_state")
        return tuple(out)

    # ------------------------------------------------------------- evaluation
    def _fock_ok(self, F):
        if F.shape != (self.n, self.n):
            raise ValueError(f"F has shape {F.shape}, expected ({self.n}, {self.n})")

    def rhs(self, kind, F, t1, t2, lam1=None, lam2=None):
        F_, pF = _c(F)
        a, pa = _c(t1)
        b, pb = _c(t2)
        self._shape_ok(a, b)
        self._fock_ok(F_)
        c, pc = _opt(lam1)
        d, pd = _opt(lam2)
        if kind in (RHS_L1, RHS_L2, RHS_L1_CC2, RHS_L2_CC2):
            if c is None or d is None:
                raise ValueError("the Lambda residuals need lam1 and lam2")
            self._shape_ok(c, d)
        elif (c is None) != (d is None) or (c is not None and (c.shape, d.shape) != (a.shape, b.shape)):
            raise ValueError("lam1/lam2 must both be given with the shapes of t1/t2, or both be None")
        one_body = kind in (RHS_T1, RHS_L1, RHS_L1_CC2)
        shape = (self.o, self.v) if one_body else (self.o, self.o, self.v, self.v)
        out = np.empty(shape, complex)
        self._check(self.lib.tdcc_rhs(self._h, int(kind), pF, pa, pb, pc, pd, out.ctypes.data_as(_dp)),
                    "tdcc_rhs")
        return out

    def intermediate(self, name, shape, F, t1, t2, lam1, lam2):
        F_, pF = _c(F)
        a, pa = _c(t1)
        b, pb = _c(t2)
        c, pc = _c(lam1)
        d, pd = _c(lam2)
        self._fock_ok(F_)
        self._shape_ok(a, b)
        self._shape_ok(c, d)
        out = np.empty(shape, complex)
        self._check(self.lib.tdcc_intermediate(self._h, name.encode(), pF, pa, pb, pc, pd,
                                               out.ctypes.data_as(_dp), out.size), f"tdcc_intermediate({name})")
        return out

    def dipole(self, t1, t2, lam1, lam2):
        a, pa = _c(t1)
        b, pb = _c(t2)
        c, pc = _c(lam1)
        d, pd = _c(lam2)
        self._shape_ok(a, b)
        self._shape_ok(c, d)
        out = np.empty(3, complex)
        self._check(self.lib.tdcc_dipole(self._h, pa, pb, pc, pd, out.ctypes.data_as(_dp)), "tdcc_dipole")
        return out

    def dipole_device_noalloc(self, t1, t2, lam1, lam2, out):
        """tdcc_dipole_device_noalloc on device memory: the amplitudes as
        interleaved complex128 in C order with the shapes (o,v), (o,o,v,v),
        (o,v), (o,o,v,v), and `out` 6 doubles (x, y, z as re, im pairs).  Each
        argument is a device address (int) or a tensor with data_ptr() (then
        contiguous complex128 of that shape, or 6 float64 for `out`).  A torch
        tensor must come from the same HIP runtime as this library: import
        torch before the library is first loaded (torch bundles its own
        runtime, which is loaded as a second copy otherwise)."""
        o, v = self.o, self.v

        def addr(x, shape, dtype_ok):
            if isinstance(x, int):
                return ctypes.c_void_p(x)
            if tuple(x.shape) != shape or not x.is_contiguous() or not dtype_ok(x):
                raise ValueError(f"expected a contiguous device tensor of shape {shape}")
            return ctypes.c_void_p(x.data_ptr())

        cplx = lambda x: str(x.dtype).endswith("complex128")
        f64 = lambda x: str(x.dtype).endswith("float64")
        ptrs = [addr(x, sh, cplx) for x, sh in ((t1, (o, v)), (t2, (o, o, v, v)), (lam1, (o, v)), (lam2, (o, o, v, v)))]
        po = addr(out, (6,), f64)
        self._check(self.lib.tdcc_dipole_device_noalloc(self._h, *ptrs, po), "tdcc_dipole_device_noalloc")
        return out

    def rk4_run(self, method, nsteps, t0, dt, w0, A, pol=2, precs=15):
        m = METHODS[method] if isinstance(method, str) else int(method)
        mu = np.empty((max(nsteps, 1), 3), complex)
        ts = np.empty(max(nsteps, 1))
        self._check(self.lib.tdcc_rk4_run(self._h, m, int(nsteps), float(t0), float(dt), float(w0), float(A),
                                          int(pol), int(precs), mu.ctypes.data_as(_dp),
                                          ts.ctypes.data_as(_dp)), "tdcc_rk4_run")
        return mu[:nsteps], ts[:nsteps]

    def profile_step(self, method, t0, dt, w0, A, pol=2, precs=15):
        """One eager step with per-launch HIP event timing: list of (name, flops, bytes, ms, detail)."""
        m = METHODS[method] if isinstance(method, str) else int(method)
        buf = ctypes.create_string_buffer(1 << 22)
        self._check(self.lib.tdcc_profile_step(self._h, m, float(t0), float(dt), float(w0), float(A), int(pol),
                                               int(precs), buf, len(buf)), "tdcc_profile_step")
        out = []
        for line in buf.value.decode().splitlines():
            name, fl, by, ms, detail = line.split("\t")
            out.append((name, float(fl), float(by), float(ms), detail))
        return out

    # ------------------------------------------------------------- batches
    def set_batch(self, nb):
        self._check(self.lib.tdcc_set_batch(self._h, int(nb)), "tdcc_set_batch")
        self.nb = int(nb)

    def set_state_slot(self, slot, t1, t2, lam1, lam2):
        a, pa = _c(t1)
        b, pb = _c(t2)
        c, pc = _c(lam1)
        d, pd = _c(lam2)
        self._shape_ok(a, b)
        self._shape_ok(c, d)
        self._check(self.lib.tdcc_set_state_slot(self._h, int(slot), pa, pb, pc, pd), "tdcc_set_state_slot")

    def get_state_slot(self, slot, which=0):
        o, v = self.o, self.v
        out = [np.empty((o, v), complex), np.empty((o, o, v, v), complex),
               np.empty((o, v), complex), np.empty((o, o, v, v), complex)]
        self._check(self.lib.tdcc_get_state_slot(self._h, int(slot), int(which),
                                                 *[x.ctypes.data_as(_dp) for x in out]), "tdcc_get_state_slot")
        return out

    def rk4_run_batch(self, method, nsteps, t0, dt, pol, w0, A, precs=15):
        """Advance every batch slot by nsteps steps; slot b uses pol[b], w0[b], A[b].
        Returns mu (nb, nsteps, 3) complex and the nsteps times."""
        m = METHODS[method] if isinstance(method, str) else int(method)
        nb = self.nb
        pol_ = (ctypes.c_int * nb)(*[int(x) for x in pol])
        w0_, pw0 = _r(np.asarray(w0, dtype=float).reshape(nb))
        A_, pA = _r(np.asarray(A, dtype=float).reshape(nb))
        mu = np.empty((nb, max(nsteps, 1), 3), complex)
        ts = np.empty(max(nsteps, 1))
        self._check(self.lib.tdcc_rk4_run_batch(self._h, m, int(nsteps), float(t0), float(dt), pol_, pw0, pA,
                                                int(precs), mu.ctypes.data_as(_dp), ts.ctypes.data_as(_dp)),
                    "tdcc_rk4_run_batch")
        return mu[:, :nsteps], ts[:nsteps]

    # ----------------------------------------------------------- gradients
    def dipole_gradient(self, method, nsteps, t0, dt, w0, A, weights, pol=2, precs=15, checkpoint_every=50):
        """L = sum_n weights[n] mu_pol(t_n) over nsteps steps from the current
        state, and dL/dA, dL/dw0 (complex scalars), dL/dmu_pol (n, n complex).
        The state is left at the end of the propagation."""
        m = METHODS[method] if isinstance(method, str) else int(method)
        w_, pw = _r(np.asarray(weights, dtype=float).reshape(int(nsteps)))
        loss = np.empty(1, complex)
        dA = np.empty(1, complex)
        dw = np.empty(1, complex)
        dmu = np.empty((self.n, self.n), complex)
        self._check(self.lib.tdcc_dipole_gradient(self._h, m, int(nsteps), float(t0), float(dt), float(w0), float(A),
                                                  int(pol), int(precs), pw, int(checkpoint_every),
                                                  loss.ctypes.data_as(_dp), dA.ctypes.data_as(_dp),
                                                  dw.ctypes.data_as(_dp), dmu.ctypes.data_as(_dp)),
                    "tdcc_dipole_gradient")
        return loss[0], dA[0], dw[0], dmu

    def checkpoint(self, op):
        self._check(self.lib.tdcc_checkpoint(self._h, int(op)), "tdcc_checkpoint")

    def step_info(self, method, dt):
        m = METHODS[method] if isinstance(method, str) else int(method)
        f = ctypes.c_int64()
        a = ctypes.c_int64()
        self._check(self.lib.tdcc_step_info(self._h, m, float(dt), ctypes.byref(f), ctypes.byref(a)),
                    "tdcc_step_info")
        return f.value, a.value


def write_restart_block(x, path, precs):
    """One restart block file (rank 2 or 4, real) through the C++ writer."""
    lib = load()
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    if a.ndim not in (2, 4):
        raise ValueError("restart blocks are rank 2 or 4")
    dims = (ctypes.c_int64 * a.ndim)(*a.shape)
    rc = lib.tdcc_write_restart_block(a.ctypes.data_as(_dp), a.ndim, dims, int(precs), os.fsencode(path))
    if rc:
        raise TdccError(f"tdcc_write_restart_block: {lib.tdcc_io_error().decode()}")


def read_restart_block(path, shape):
    """One restart block file (rank 2 or 4) through the C++ reader: a dense
    float64 array of `shape` (zero where the file has no row)."""
    lib = load()
    shape = tuple(int(x) for x in shape)
    if len(shape) not in (2, 4):
        raise ValueError("restart blocks are rank 2 or 4")
    out = np.empty(shape, dtype=np.float64)
    dims = (ctypes.c_int64 * len(shape))(*shape)
    rc = lib.tdcc_read_restart_block(os.fsencode(path), len(shape), dims, out.ctypes.data_as(_dp))
    if rc:
        raise TdccError(f"tdcc_read_restart_block: {lib.tdcc_io_error().decode()}")
    return out


def write_state_file(path, method, precs, t, dt, w0, A, F, t1, t2, lam1, lam2):
    """Exact binary snapshot of a stopped run (include/tdcc.h, tdcc_write_state_file)."""
    lib = load()
    o, v = np.shape(t1)
    F_, pF = _r(np.real(F))
    arrs = [_c(x) for x in (t1, t2, lam1, lam2)]
    if F_.shape != (o + v, o + v) or arrs[1][0].shape != (o, o, v, v) or arrs[2][0].shape != (o, v) \
            or arrs[3][0].shape != (o, o, v, v):
        raise ValueError("state file arrays do not have the shapes of one (o, v) system")
    params = np.array([t, dt, w0, A], dtype=np.float64)
    rc = lib.tdcc_write_state_file(os.fsencode(path), o, v, METHODS[method], int(precs), params.ctypes.data_as(_dp),
                                   pF, *[p for _, p in arrs])
    if rc:
        raise TdccError(f"tdcc_write_state_file: {lib.tdcc_io_error().decode()}")


def read_state_file(path):
    """Everything tdcc_write_state_file stored: dict with method, precs, t, dt,
    w0, A, F (real), t1, t2, lam1, lam2."""
    lib = load()
    o, v, m, pr = (ctypes.c_int() for _ in range(4))
    params = np.empty(4)
    if lib.tdcc_read_state_file_header(os.fsencode(path), ctypes.byref(o), ctypes.byref(v), ctypes.byref(m),
                                       ctypes.byref(pr), params.ctypes.data_as(_dp)):
        raise TdccError(f"tdcc_read_state_file_header: {lib.tdcc_io_error().decode()}")
    o, v = o.value, v.value
    F = np.empty((o + v, o + v))
    out = [np.empty((o, v), complex), np.empty((o, o, v, v), complex), np.empty((o, v), complex),
           np.empty((o, o, v, v), complex)]
    if lib.tdcc_read_state_file(os.fsencode(path), F.ctypes.data_as(_dp), *[x.ctypes.data_as(_dp) for x in out]):
        raise TdccError(f"tdcc_read_state_file: {lib.tdcc_io_error().decode()}")
    method = {v_: k for k, v_ in METHODS.items()}[m.value]
    return dict(method=method, precs=pr.value, t=float(params[0]), dt=float(params[1]), w0=float(params[2]),
                A=float(params[3]), F=F, t1=out[0], t2=out[1], lam1=out[2], lam2=out[3])
