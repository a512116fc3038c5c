"""Drop-in CCSD_Helper for the real-time propagation path, backed by the HIP library.

Mirrors the reference class (CCSD_Helper.py:24-1256) on the methods the
propagation path uses, with the same names, argument order and return values:
the T/Lambda right-hand sides and their intermediates, the density blocks and
dipole, the restart writers and Runge_Kutta_solver.  Every evaluation runs on
the GPU through libtdcc_hip.so (include/tdcc.h); there is no CPU path.

The constructor takes a system provider in place of psi4 (the reference
ignores its `psi` argument and reads the module-level psi4, CCSD_Helper.py:26-35).
A provider has `nocc`, `nvir`, `F` (n,n), `mu` (3,n,n) and either `TEI`
(n,n,n,n) or `tei_config = (config_id, sigma)` for device-side generation;
`python_code_to_check_c_amd.synthetic.SyntheticSystem` is one.
"""
from __future__ import annotations

import time

import numpy as np
import pandas as pd

from . import _lib
from . import restart_io


class CCSD_Helper(object):
    method = "CCSD"
    time_digits = 4

    def __init__(self, psi, device=0, digits=0):
        """`digits` > 0 runs the large contractions in split precision with that
        many int8 digits per operand (include/tdcc.h, tdcc_set_precision); the
        default 0 keeps the reference's FP64 arithmetic throughout."""
        self.system = psi
        o, v = int(psi.nocc), int(psi.nvir)
        self.nocc, self.nvir = o, v
        n = o + v
        self.nmo = n // 2
        self.ndocc = o // 2
        self.occ = slice(0, o)
        self.vir = slice(o, n)
        self.TEI = getattr(psi, "TEI", None)
        self.counter = 0
        self.verbose = True
        self.ctx = _lib.Context(o, v, device)
        self.ctx.set_system(self.TEI, psi.F, psi.mu)
        if self.TEI is None:
            cid, sigma = psi.tei_config
            self.ctx.synth_tei(cid, sigma)
        self._F_loaded = np.asarray(psi.F, dtype=float)
        self.digits = int(digits)
        self.ctx.set_precision(self.digits)

    # ------------------------------------------------------------ system
    def F_MO(self, H=None, C=None):
        """Spin-orbital Fock matrix (CCSD_Helper.py:82-93)."""
        return np.array(self.system.F, dtype=float)

    def Defd_dipole(self):
        """Spin-orbital dipole matrices x, y, z (CCSD_Helper.py:970-980)."""
        mu = np.asarray(self.system.mu)
        return [mu[0], mu[1], mu[2]]

    def _use_fock(self, F):
        F = np.asarray(F)
        if F.dtype.kind == "c" or F.shape != self._F_loaded.shape or not np.array_equal(F, self._F_loaded):
            Fr = np.ascontiguousarray(F.real, dtype=float)
            self.ctx.set_system(None, Fr, self.system.mu)
            self._F_loaded = Fr

    def _shape(self, name):
        o, v = self.nocc, self.nvir
        shapes = {"Fae": (v, v), "Fmi": (o, o), "Fme": (o, v), "Wmnij": (o, o, o, o), "Wmnij_2": (o, o, o, o),
                  "Wabef": (v, v, v, v), "Wabef_2": (v, v, v, v), "Wmbej": (o, v, v, o), "LRFea": (v, v),
                  "LRFim": (o, o), "LSWieam": (o, v, v, o), "LRWefam": (v, v, v, o), "LRWibjm": (o, v, o, o),
                  "LSWmnij": (o, o, o, o), "LSWabef": (v, v, v, v), "Gfe": (v, v), "Gmn": (o, o),
                  "LWfiea": (v, o, v, v), "LWmina": (o, o, o, v), "Dij": (o, o), "Dab": (v, v), "Dai": (v, o)}
        return shapes[name]

    def _inter(self, name, t1, t2, F=None, lam1=None, lam2=None):
        o, v = self.nocc, self.nvir
        n = o + v
        F = np.zeros((n, n)) if F is None else F
        lam1 = np.zeros((o, v), complex) if lam1 is None else lam1
        lam2 = np.zeros((o, o, v, v), complex) if lam2 is None else lam2
        t1 = np.zeros((o, v), complex) if t1 is None else t1
        t2 = np.zeros((o, o, v, v), complex) if t2 is None else t2
        return self.ctx.intermediate(name, self._shape(name), F, t1, t2, lam1, lam2)

    # -------------------------------------------------- T intermediates / RHS
    def Fae(self, t1, t2, F):
        return self._inter("Fae", t1, t2, F)

    def Fmi(self, t1, t2, F):
        return self._inter("Fmi", t1, t2, F)

    def Fme(self, t1, t2, F):
        return self._inter("Fme", t1, t2, F)

    def Wmnij(self, t1, t2, F):
        return self._inter("Wmnij", t1, t2, F)

    def Wmnij_2(self, t1, t2, F):
        return self._inter("Wmnij_2", t1, t2, F)

    def Wabef(self, t1, t2, F):
        return self._inter("Wabef", t1, t2, F)

    def Wabef_2(self, t1, t2, F):
        return self._inter("Wabef_2", t1, t2, F)

    def Wmbej(self, t1, t2, F):
        return self._inter("Wmbej", t1, t2, F)

    def T1eq_rhs(self, t1, t2, F):
        """CCSD_Helper.py:177-196"""
        return self.ctx.rhs(_lib.RHS_T1, F, t1, t2)

    def T2eq_rhs(self, t1, t2, F):
        """CCSD_Helper.py:268-309"""
        return self.ctx.rhs(_lib.RHS_T2, F, t1, t2)

    def CCSD_Corr_E(self, t1, t2, F):
        """CCSD_Helper.py:312-321 (O(o^2 v^2) host reduction)."""
        o, v = self.occ, self.vir
        oovv = self._oovv()
        return (np.einsum("ia,ia->", F[o, v], t1) + 0.25 * np.einsum("ijab,ijab->", oovv, t2)
                + 0.5 * np.einsum("ijab,ia,jb->", oovv, t1, t1))

    def _oovv(self):
        if self.TEI is not None:
            return self.TEI[self.occ, self.occ, self.vir, self.vir]
        return self.system.tei_block("oovv")

    # --------------------------------------------- Lambda intermediates / RHS
    def LRFea(self, t1, t2, F):
        return self._inter("LRFea", t1, t2, F)

    def LRFim(self, t1, t2, F):
        return self._inter("LRFim", t1, t2, F)

    def LSWieam(self, t1, t2, F):
        return self._inter("LSWieam", t1, t2, F)

    def LRWefam(self, t1, t2, F):
        return self._inter("LRWefam", t1, t2, F)

    def LRWibjm(self, t1, t2, F):
        return self._inter("LRWibjm", t1, t2, F)

    def Gfe(self, t2, lam2):
        return self._inter("Gfe", None, t2, lam2=lam2)

    def Gmn(self, t2, lam2):
        return self._inter("Gmn", None, t2, lam2=lam2)

    def LWfiea(self, t1):
        return self._inter("LWfiea", t1, None)

    def LWmina(self, t1):
        return self._inter("LWmina", t1, None)

    def LSWmnij(self, t1, t2, F):
        return self._inter("LSWmnij", t1, t2, F)

    def LSWabef(self, t1, t2, F):
        return self._inter("LSWabef", t1, t2, F)

    def lam_1eq_rhs(self, t1, t2, lam1, lam2, F):
        """CCSD_Helper.py:555-581"""
        return self.ctx.rhs(_lib.RHS_L1, F, t1, t2, lam1, lam2)

    def lam2eq_rhs(self, t1, t2, lam1, lam2, F):
        """CCSD_Helper.py:610-648"""
        return self.ctx.rhs(_lib.RHS_L2, F, t1, t2, lam1, lam2)

    def CCSD_pseudo_E(self, t1, t2, lam1, lam2, F):
        """CCSD_Helper.py:650-656"""
        return (np.einsum("ia,ia->", F[self.occ, self.vir], lam1),
                0.25 * np.einsum("ijab,ijab->", self._oovv(), lam2))

    # ------------------------------------------------------ density / dipole
    def Dij(self, t1, t2, lam1, lam2):
        return self._inter("Dij", t1, t2, None, lam1, lam2)

    def Dab(self, t1, t2, lam1, lam2):
        return self._inter("Dab", t1, t2, None, lam1, lam2)

    def Dai(self, t1, t2, lam1, lam2):
        return self._inter("Dai", t1, t2, None, lam1, lam2)

    def dipole_moment(self, t1, t2, lam1, lam2, F):
        """CCSD_Helper.py:1066-1092: [mu_x, mu_y, mu_z] of the correlated density."""
        return list(self.ctx.dipole(t1, t2, lam1, lam2))

    # --------------------------------------------------------------- output
    def Save_parameters(self, w0, A, t0, t, dt, precs, i, a, CCSD_or_CC2):
        restart_io.save_parameters(w0, A, t, dt, precs, i, a, CCSD_or_CC2)

    def write_2data(self, F, FileName, precs):
        restart_io.write_2data(F, FileName, precs)

    def write_4data(self, F, FileName, precs):
        restart_io.write_4data(F, FileName, precs)

    def Save_data(self, F, t1, t2, lam1, lam2, data, timing, precs, restart):
        restart_io.save_data(F, t1, t2, lam1, lam2, data, timing, precs, restart)

    # ------------------------------------------------------------ propagator
    def Runge_Kutta_solver(self, F, t1, t2, lam1, lam2, w0, A, t0, tf, dt, timeout, precs, restart=None):
        """CCSD_Helper.py:1191-1256."""
        return self._propagate(F, t1, t2, lam1, lam2, w0, A, t0, tf, dt, timeout, precs, restart)

    def _propagate(self, F, t1, t2, lam1, lam2, w0, A, t0, tf, dt, timeout, precs, restart,
                   pol=2, max_steps=None, chunk_seconds=0.5, max_chunk=256):
        """The reference loop on the device, in chunks of graph-replayed steps.

        On a stop the reference's text files are written (rounded base of the
        stopping step, time t - dt) and, beside them, the exact snapshot of the
        state after that step and its time t (restart_io.save_state), from
        which CC_Calculator.TDCC_restart(exact=True) resumes bit for bit.

        The stop rules are applied after every step, in the reference's order
        (CCSD_Helper.py:1234-1251): first the wall clock (elapsed time at the
        end of the step's update above 0.9 * timeout minutes), then divergence
        (|Re mu_z| > 100).  Inside a chunk the end of step j is taken as the
        chunk's start plus (j + 1) times its mean step time.  When a rule fires
        before the chunk's last step, the state is restored to the chunk start
        and the chunk is replayed up to that step (the kernels are
        deterministic), so the saved rounded base and time are those of the
        stopping step.  Running to t >= tf saves nothing.

        Timing rows: `total` is the mean step time of the step's chunk; the
        per-chain columns t1, t2, l1, l2 are NaN, because one captured graph
        runs the four chains together, sharing the intermediates the
        reference recomputes per chain, so there is no per-chain time to take.
        """
        ctx = self.ctx
        method = self.method
        self._use_fock(F)
        ctx.set_state(t1, t2, lam1, lam2)
        t = t0
        rows, trows = [], []
        start = time.time()
        mua = self.dipole_moment(t1, t2, lam1, lam2, F)
        if self.verbose:
            print("time \t\t mu_z_real \t\t mu_z_real")
            print(round(t, 4), '\t', mua[2].real, '\t', mua[2].imag)
        i = 0
        chunk = 1
        saved = None
        mu_all = []
        while t < tf and (max_steps is None or i < max_steps):
            # number of steps before t reaches tf, with the reference's float accumulation
            k, tt = 0, t
            while tt < tf and k < chunk and (max_steps is None or i + k < max_steps):
                tt += dt
                k += 1
            ctx.checkpoint(0)
            c0 = time.time()
            mu, ts = ctx.rk4_run(method, k, t, dt, w0, A, pol, precs)
            per = (time.time() - c0) / k
            stop_at, reason = None, None
            for j in range(k):
                if abs(c0 + (j + 1) * per - start) > 0.9 * timeout * 60.0:
                    stop_at, reason = j, "timeout"
                elif abs(mu[j, 2].real) > 100:
                    stop_at, reason = j, "divergence"
                if stop_at is not None:
                    break
            if stop_at is not None and stop_at < k - 1:
                ctx.checkpoint(1)
                mu, ts = ctx.rk4_run(method, stop_at + 1, t, dt, w0, A, pol, precs)
                k = stop_at + 1
            for j in range(k):
                i += 1
                t = float(ts[j])
                rows.append((i, round(t, self.time_digits), mu[j, 2].real, mu[j, 2].imag))
                trows.append((i, per, np.nan, np.nan, np.nan, np.nan))
                mu_all.append(mu[j])
                if self.verbose:
                    print(round(t, self.time_digits), '\t', mu[j, 2].real, '\t', mu[j, 2].imag)
            if stop_at is not None:
                base = ctx.get_state(1)
                data, timing = self._frames(rows, trows)
                self.Save_data(F, *base, data, timing, precs, restart)
                self.Save_parameters(w0, A, t0, t - dt, dt, precs, self.nocc, self.nvir, method)
                restart_io.save_state(F, *ctx.get_state(0), t, dt, w0, A, precs, method)
                saved = reason
                break
            if per * chunk < chunk_seconds:
                chunk = min(max_chunk, chunk * 2)
        stop = time.time()
        if self.verbose and i:
            print("total time non-adapative step:", stop - start)
            print("total steps:", i)
            print("step-time:", (stop - start) / i)
        self.data, self.timing = self._frames(rows, trows)
        self.mu_history = np.array(mu_all).reshape(-1, 3)
        self.stop_reason = saved
        self.final_time = t
        return None

    @staticmethod
    def _frames(rows, trows):
        data = pd.DataFrame([r[1:] for r in rows], index=[r[0] for r in rows],
                            columns=("time", "mu_real", "mu_imag"))
        timing = pd.DataFrame([r[1:] for r in trows], index=[r[0] for r in trows],
                              columns=("total", "t1", "t2", "l1", "l2"))
        return data, timing

    def get_state(self):
        """Current device state (t1, t2, lam1, lam2) after a propagation."""
        return self.ctx.get_state(0)
