"""numpy complex128 restatement of the reference's coupled-cluster equations.

Test infrastructure only (see oracle/__init__.py).

Every method here restates one reference function, keeps its name and its
argument order, and cites the reference lines it follows.  The einsum strings
and coefficients are the reference's own (an einsum output label is only a
position, so strings are kept exactly, never relabelled).  `contract` is
`numpy.einsum(..., optimize='optimal')`, which replaces the reference's
opt_einsum.contract (unpinned third-party dependency, no version recorded in the
reference; only the summation order differs).

Conventions: TEI is the real antisymmetrized spin-orbital <pq||rs> (n,n,n,n),
occupied orbitals first; t1/lam1 are (o,v), t2/lam2 (o,o,v,v); F is (n,n) and
may be complex (F + V(t) inside the propagator).
"""
from __future__ import annotations

import numpy as np


def np_contract(spec, *ops):
    return np.einsum(spec, *ops, optimize="optimal")


def _swap_ij(x):
    return x.swapaxes(0, 1)


def _swap_ab(x):
    return x.swapaxes(2, 3)


def _anti01(x):
    """x - x with positions 0,1 exchanged."""
    return x - x.swapaxes(0, 1)


def _anti23(x):
    """x - x with positions 2,3 exchanged."""
    return x - x.swapaxes(2, 3)


def _anti_both(x):
    """x - x(2<->3) - x(0<->1) + x(0<->1, 2<->3), in the reference's order."""
    return x - x.swapaxes(2, 3) - x.swapaxes(0, 1) + x.swapaxes(0, 1).swapaxes(2, 3)


class CCEquations:
    """State-free CCSD/CC2 equations over one set of integrals."""

    def __init__(self, TEI, nocc, contract=None, vchunk=None):
        """vchunk: None evaluates every v^4 intermediate (Wabef, Wabef_2,
        LSWabef, the <ab||ef> t1 term of L1eq_rhs_CC2) whole, as the
        reference does; an integer builds it `vchunk` leading virtual indices
        at a time and contracts each slab at once (the same einsum strings on
        index slices), so that o=20, v=160 fits in host memory."""
        self.TEI = np.asarray(TEI, dtype=np.float64)
        n = self.TEI.shape[0]
        self.nmo = n // 2
        self.nocc = nocc
        self.nvir = n - nocc
        self.occ = slice(0, nocc)
        self.vir = slice(nocc, n)
        self.contract = contract or np_contract
        self.vchunk = vchunk

    def _chunks(self):
        step = self.vchunk or self.nvir
        return [slice(a, min(a + step, self.nvir)) for a in range(0, self.nvir, step)]

    def _sub(self, spec, ops, letter, rows):
        """contract(spec, *ops) with the output letter `letter` restricted to
        `rows` (every operand axis carrying that letter is sliced)."""
        ins = spec.split("->")[0].split(",")
        cut = [op[tuple(rows if ch == letter else slice(None) for ch in lab)] for lab, op in zip(ins, ops)]
        return self.contract(spec, *cut)

    def _rows_anti01(self, spec, ops, rows):
        """Rows `rows` of _anti01(contract(spec, *ops)) for a rank-4 output."""
        ins, out = spec.split("->")
        swapped = ins + "->" + out[1] + out[0] + out[2:]
        return self._sub(spec, ops, out[0], rows) - self._sub(swapped, ops, out[1], rows)

    def blk(self, key):
        """TEI block by orbital classes, e.g. 'oovv' -> TEI[o, o, v, v]."""
        sl = {"o": self.occ, "v": self.vir}
        return self.TEI[tuple(sl[c] for c in key)]

    def _t1t1(self, t1):
        return self.contract("ia,jb->ijab", t1, t1)

    # ---------------------------------------------------------------- T side
    def Fae(self, t1, t2, F):
        """CCSD_Helper.py:136-147"""
        c, o, v = self.contract, self.occ, self.vir
        tau = t2.copy() + self._t1t1(t1)
        return (F[v, v].copy()
                - 0.5 * c("me,ma->ae", F[o, v], t1)
                + c("mafe,mf->ae", self.blk("ovvv"), t1)
                - 0.5 * c("mnef,mnaf->ae", self.blk("oovv"), tau))

    def Fmi(self, t1, t2, F):
        """CCSD_Helper.py:150-160"""
        c, o = self.contract, self.occ
        tau = t2.copy() + self._t1t1(t1)
        return (F[o, o].copy()
                + 0.5 * c("me,ie->mi", F[o, self.vir], t1)
                + c("mnie,ne->mi", self.blk("ooov"), t1)
                + 0.5 * c("mnef,inef->mi", self.blk("oovv"), tau))

    def Fme(self, t1, t2, F):
        """CCSD_Helper.py:163-170"""
        return F[self.occ, self.vir].copy() + self.contract("mnef,nf->me", self.blk("oovv"), t1)

    def T1eq_rhs(self, t1, t2, F):
        """CCSD_Helper.py:177-196"""
        c = self.contract
        fae = self.Fae(t1, t2, F)
        fmi = self.Fmi(t1, t2, F)
        fme = self.Fme(t1, t2, F)
        return (F[self.occ, self.vir].copy()
                + c("ae,ie->ia", fae, t1)
                - c("mi,ma->ia", fmi, t1)
                + c("me,imae->ia", fme, t2)
                - c("naif,nf->ia", self.blk("ovov"), t1)
                - 0.5 * c("nmei,mnae->ia", self.blk("oovo"), t2)
                - 0.5 * c("maef,imef->ia", self.blk("ovvv"), t2))

    def Wmnij(self, t1, t2, F):
        """CCSD_Helper.py:199-210"""
        c = self.contract
        tau = (0.25 * t2 + 0.5 * c("ia,jb->ijab", t1, t1)
               - 0.5 * c("ib,ja->ijab", t1, t1))
        return (self.blk("oooo").copy()
                + _anti23(c("mnie,je->mnij", self.blk("ooov"), t1))
                + c("mnef,ijef->mnij", self.blk("oovv"), tau))

    def Wmnij_2(self, t1, t2, F):
        """CCSD_Helper.py:213-224"""
        c = self.contract
        tau = self._t1t1(t1)
        return (self.blk("oooo").copy()
                + _anti23(c("mnie,je->mnij", self.blk("ooov"), t1))
                + _anti23(0.25 * c("mnef,ijef->mnij", self.blk("oovv"), tau)))

    def Wabef(self, t1, t2, F, rows=None):
        """CCSD_Helper.py:227-239 (rows: the slab a in rows only)"""
        c = self.contract
        tau = self._t1t1(t1)
        if rows is not None:
            vovv, oovv = self.blk("vovv"), self.blk("oovv")
            return (self.blk("vvvv")[rows].copy()
                    + self._rows_anti01("amef,mb->abef", (vovv, -t1), rows)
                    + 0.25 * self._sub("mnef,mnab->abef", (oovv, t2), "a", rows)
                    + self._rows_anti01("mnef,mnab->abef", (oovv, 0.5 * tau), rows))
        return (self.blk("vvvv").copy()
                + _anti01(-c("amef,mb->abef", self.blk("vovv"), t1))
                + 0.25 * c("mnef,mnab->abef", self.blk("oovv"), t2)
                + _anti01(0.5 * c("mnef,mnab->abef", self.blk("oovv"), tau)))

    def Wabef_2(self, t1, t2, F, rows=None):
        """CCSD_Helper.py:242-252 (rows: the slab a in rows only)"""
        c = self.contract
        tau = self._t1t1(t1) - c("ib,ja->ijab", t1, t1)
        if rows is not None:
            vovv, oovv = self.blk("vovv"), self.blk("oovv")
            return (self.blk("vvvv")[rows].copy()
                    + self._rows_anti01("amef,mb->abef", (vovv, -t1), rows)
                    + 0.25 * self._sub("mnef,mnab->abef", (oovv, tau), "a", rows))
        return (self.blk("vvvv").copy()
                + _anti01(-c("amef,mb->abef", self.blk("vovv"), t1))
                + 0.25 * c("mnef,mnab->abef", self.blk("oovv"), tau))

    def Wmbej(self, t1, t2, F):
        """CCSD_Helper.py:255-265"""
        c = self.contract
        t2t1 = 0.5 * t2 + c("jf,nb->jnfb", t1, t1)
        return (self.blk("ovvo").copy()
                - c("mnej,nb->mbej", self.blk("oovo"), t1)
                - c("mnef,jnfb->mbej", self.blk("oovv"), t2t1)
                + c("mbef,jf->mbej", self.blk("ovvv"), t1))

    def T2eq_rhs(self, t1, t2, F):
        """CCSD_Helper.py:268-309"""
        c = self.contract
        fae = self.Fae(t1, t2, F)
        fmi = self.Fmi(t1, t2, F)
        fme = self.Fme(t1, t2, F)
        wmnij = self.Wmnij(t1, t2, F)
        wmbej = self.Wmbej(t1, t2, F)
        wmnij_2 = self.Wmnij_2(t1, t2, F)
        fvv = fae - 0.5 * c("me,mb->be", fme, t1)
        foo = fmi + 0.5 * c("me,je->mj", fme, t1)
        tau = c("ma,nb->mnab", t1, t1) - c("na,mb->mnab", t1, t1)
        oooo_ladder = (0.5 * c("mnij,mnab->ijab", wmnij, t2)
                       + 0.5 * c("mnij,mnab->ijab", wmnij_2, tau))
        if self.vchunk is None:
            wabef = self.Wabef(t1, t2, F)
            wabef_2 = self.Wabef_2(t1, t2, F)
            vvvv_ladder = (0.5 * c("abef,ijef->ijab", wabef, t2)
                           + 0.5 * c("abef,ijef->ijab", wabef_2, tau))
        else:
            vvvv_ladder = np.empty_like(t2)
            for A in self._chunks():
                vvvv_ladder[:, :, A] = (0.5 * c("abef,ijef->ijab", self.Wabef(t1, t2, F, A), t2)
                                        + 0.5 * c("abef,ijef->ijab", self.Wabef_2(t1, t2, F, A), tau))
        ring = (c("mbej,imae->ijab", wmbej, t2)
                - c("mbej,ie,ma->ijab", self.blk("ovvo"), t1, t1))
        return (self.blk("oovv").copy()
                + _anti23(c("be,ijae->ijab", fvv, t2))
                + _anti01(-c("mj,imab->ijab", foo, t2))
                + oooo_ladder
                + vvvv_ladder
                + _anti_both(ring)
                + _anti01(c("abej,ie->ijab", self.blk("vvvo"), t1))
                + _anti23(-c("mbij,ma->ijab", self.blk("ovoo"), t1)))

    def CCSD_Corr_E(self, t1, t2, F):
        """CCSD_Helper.py:312-321"""
        c, o, v = self.contract, self.occ, self.vir
        return (c("ia,ia->", F[o, v], t1)
                + 0.25 * c("ijab,ijab->", self.blk("oovv"), t2)
                + 0.5 * c("ijab,ia,jb->", self.blk("oovv"), t1, t1))

    # ----------------------------------------------------------- Lambda side
    def LRFea(self, t1, t2, F):
        """CCSD_Helper.py:447-459"""
        c, o, v = self.contract, self.occ, self.vir
        tau = 0.5 * t2 + self._t1t1(t1)
        return (F[v, v].copy()
                - c("ma,me->ea", F[o, v], t1)
                + c("emaf,mf->ea", self.blk("vovv"), t1)
                - c("mnaf,mnef->ea", self.blk("oovv"), tau))

    def LRFim(self, t1, t2, F):
        """CCSD_Helper.py:462-472"""
        c, o, v = self.contract, self.occ, self.vir
        tau = 0.5 * t2 + self._t1t1(t1)
        return (F[o, o].copy()
                + c("ie,me->im", F[o, v], t1)
                + c("inmf,nf->im", self.blk("ooov"), t1)
                + c("inef,mnef->im", self.blk("oovv"), tau))

    def LSWieam(self, t1, t2, F):
        """CCSD_Helper.py:475-485"""
        c = self.contract
        tau = t2 + self._t1t1(t1)
        return (self.blk("ovvo").copy()
                + c("eifa,mf->ieam", self.blk("vovv"), t1)
                - c("nima,ne->ieam", self.blk("ooov"), t1)
                - c("ijab,mjbe->ieam", self.blk("oovv"), tau))

    def LRWefam(self, t1, t2, F):
        """CCSD_Helper.py:488-505"""
        c = self.contract
        fme = self.Fme(t1, t2, F)
        if self.vchunk is None:
            w_t1 = c("efab,mb->efam", self.LSWabef(t1, t2, F), t1)
        else:
            w_t1 = np.empty((self.nvir, self.nvir, self.nvir, self.nocc), dtype=np.result_type(t1, t2, F))
            for E in self._chunks():
                w_t1[E] = c("efab,mb->efam", self.LSWabef(t1, t2, F, E), t1)
        tau = 0.25 * t2 + 0.5 * self._t1t1(t1)
        ring_a = -self.blk("ovvo").copy() + c("jnab,nmfb->jfam", self.blk("oovv"), t2)
        return 0.5 * self.blk("vvvo").copy() + (
            0.5 * c("na,mnef->efam", fme, t2)
            + w_t1
            + c("jfam,je->efam", ring_a, t1)
            + c("jnam,jnef->efam", self.blk("oovo"), tau)
            - c("jfab,jmeb->efam", self.blk("ovvv"), t2))

    def LRWibjm(self, t1, t2, F):
        """CCSD_Helper.py:508-524"""
        c = self.contract
        fme = self.Fme(t1, t2, F)
        wmnij = self.LSWmnij(t1, t2, F)
        tau = 0.25 * t2 + 0.5 * self._t1t1(t1)
        ring_a = -self.blk("ovvo").copy() - c("inef,nmfb->ibem", self.blk("oovv"), t2)
        return -0.5 * self.blk("ovoo").copy() + (
            0.5 * c("ie,jmbe->ibjm", fme, t2)
            + c("injm,nb->ibjm", wmnij, t1)
            + c("ibem,je->ibjm", ring_a, t1)
            - c("ibef,jmef->ibjm", self.blk("ovvv"), tau)
            + c("inem,jneb->ibjm", self.blk("oovo"), t2))

    def Gfe(self, t2, lam2):
        """CCSD_Helper.py:526-527"""
        return -0.5 * self.contract("mnfb,mneb->fe", lam2, t2)

    def Gmn(self, t2, lam2):
        """CCSD_Helper.py:529-530"""
        return 0.5 * self.contract("njed,mjed->nm", lam2, t2)

    def LWfiea(self, t1):
        """CCSD_Helper.py:533-540"""
        return self.blk("vovv").copy() - self.contract("jiea,jf->fiea", self.blk("oovv"), t1)

    def LWmina(self, t1):
        """CCSD_Helper.py:543-550"""
        return self.blk("ooov").copy() + self.contract("mifa,nf->mina", self.blk("oovv"), t1)

    def lam_1eq_rhs(self, t1, t2, lam1, lam2, F):
        """CCSD_Helper.py:555-581"""
        c = self.contract
        fia = self.Fme(t1, t2, F)
        fea = self.LRFea(t1, t2, F)
        fim = self.LRFim(t1, t2, F)
        wieam = self.LSWieam(t1, t2, F)
        wefam = self.LRWefam(t1, t2, F)
        wibjm = self.LRWibjm(t1, t2, F)
        gef = self.Gfe(t2, lam2)
        gmn = self.Gmn(t2, lam2)
        weifa = self.LWfiea(t1)
        wmina = self.LWmina(t1)
        return (fia.copy()
                + (c("ea,ie->ia", fea, lam1)
                   - c("im,ma->ia", fim, lam1)
                   + c("ieam,me->ia", wieam, lam1))
                + c("efam,imef->ia", wefam, lam2)
                + c("ibjm,jmab->ia", wibjm, lam2)
                - c("fe,fiea->ia", gef, weifa)
                - c("nm,mina->ia", gmn, wmina))

    def LSWmnij(self, t1, t2, F):
        """CCSD_Helper.py:586-595"""
        c = self.contract
        tau = 0.25 * t2 + 0.5 * self._t1t1(t1)
        return (0.5 * self.blk("oooo").copy()
                + c("ijme,ne->ijmn", self.blk("ooov"), t1)
                + c("ijfe,mnfe->ijmn", self.blk("oovv"), tau))

    def LSWabef(self, t1, t2, F, rows=None):
        """CCSD_Helper.py:598-607 (rows: the slab e in rows only)"""
        c = self.contract
        tau = 0.25 * t2 + 0.5 * self._t1t1(t1)
        if rows is not None:
            return (0.5 * self.blk("vvvv")[rows].copy()
                    - self._sub("emab,mf->efab", (self.blk("vovv"), t1), "e", rows)
                    + self._sub("nmab,nmef->efab", (self.blk("oovv"), tau), "e", rows))
        return (0.5 * self.blk("vvvv").copy()
                - c("emab,mf->efab", self.blk("vovv"), t1)
                + c("nmab,nmef->efab", self.blk("oovv"), tau))

    def lam2eq_rhs(self, t1, t2, lam1, lam2, F):
        """CCSD_Helper.py:610-648"""
        c = self.contract
        feb = self.LRFea(t1, t2, F)
        fjm = self.LRFim(t1, t2, F)
        wijmn = self.LSWmnij(t1, t2, F)
        wjebm = self.LSWieam(t1, t2, F)
        wejab = self.LWfiea(t1)
        wijmb = self.LWmina(t1)
        fjb = self.Fme(t1, t2, F)
        gbe = self.Gfe(t2, lam2)
        gmj = self.Gmn(t2, lam2)
        oovv = self.blk("oovv")
        ring = c("jebm,imae->ijab", wjebm, lam2) + c("jb,ia->ijab", fjb, lam1)
        total = (oovv
                 + _anti23(c("eb,ijae->ijab", feb, lam2))
                 + _anti01(-c("jm,imab->ijab", fjm, lam2))
                 + c("ijmn,mnab->ijab", wijmn, lam2)
                 + (self._lam2_ladder(t1, t2, lam2, F)
                    + _anti01(c("ejab,ie->ijab", wejab, lam1)))
                 + _anti23(-c("ijmb,ma->ijab", wijmb, lam1))
                 + _anti_both(ring))
        return (total
                + _anti23(c("ijfb,af->ijab", oovv, gbe))
                + _anti01(-c("mjab,im->ijab", oovv, gmj)))

    def _lam2_ladder(self, t1, t2, lam2, F):
        """c("efab,ijef->ijab", LSWabef, lam2) (CCSD_Helper.py:631); chunked,
        the sum over e is taken slab by slab."""
        c = self.contract
        if self.vchunk is None:
            return c("efab,ijef->ijab", self.LSWabef(t1, t2, F), lam2)
        out = None
        for E in self._chunks():
            part = c("efab,ijef->ijab", self.LSWabef(t1, t2, F, E), lam2[:, :, E])
            out = part if out is None else out + part
        return out

    def CCSD_pseudo_E(self, t1, t2, lam1, lam2, F):
        """CCSD_Helper.py:650-656"""
        c, o, v = self.contract, self.occ, self.vir
        return (c("ia,ia->", F[o, v], lam1),
                0.25 * c("ijab,ijab->", self.blk("oovv"), lam2))

    # ------------------------------------------------------------------ CC2
    def T2eq_rhs_CC2(self, t1, t2, F):
        """CC2_Helper.py:12-47"""
        c, o, v = self.contract, self.occ, self.vir
        wmnij_2 = self.Wmnij_2(t1, t2, F)
        tau = c("ma,nb->mnab", t1, t1) - c("na,mb->mnab", t1, t1)
        if self.vchunk is None:
            ladder = 0.5 * c("abef,ijef->ijab", self.Wabef_2(t1, t2, F), tau)
        else:
            ladder = np.empty(tau.shape, dtype=np.result_type(t1, t2, F))
            for A in self._chunks():
                ladder[:, :, A] = 0.5 * c("abef,ijef->ijab", self.Wabef_2(t1, t2, F, A), tau)
        ring = -c("mbej,ie,ma->ijab", self.blk("ovvo"), t1, t1)
        return (self.blk("oovv").copy()
                + _anti23(c("be,ijae->ijab", F[v, v].copy(), t2))
                + _anti01(-c("mj,imab->ijab", F[o, o].copy(), t2))
                + 0.5 * c("mnij,mnab->ijab", wmnij_2, tau)
                + ladder
                + _anti_both(ring)
                + _anti01(c("abej,ie->ijab", self.blk("vvvo"), t1))
                + _anti23(-c("mbij,ma->ijab", self.blk("ovoo"), t1)))

    def L1eq_rhs_CC2(self, t1, t2, lam1, lam2, F):
        """CC2_Helper.py:50-125 (coefficients as written, e.g. the 1/2 at :115, :117)"""
        c = self.contract
        oovv = self.blk("oovv")
        rhs = self.Fme(t1, t2, F).copy()
        rhs = rhs + c("ea,ie->ia", self.LRFea(t1, t2, F), lam1)
        rhs = rhs - c("im,ma->ia", self.LRFim(t1, t2, F), lam1)
        t1t1 = self._t1t1(t1)
        wmbej = (self.blk("ovvo").copy()
                 + c("mbef,jf->mbej", self.blk("ovvv"), t1)
                 - c("mnej,nb->mbej", self.blk("oovo"), t1)
                 - c("mnef,jnfb->mbej", oovv, t1t1))
        gbj = -c("me,mjbe->bj", lam1, t2)
        rhs = rhs + c("ijab,bj->ia", oovv, gbj)
        rhs = rhs + c("ieam,me->ia", wmbej, lam1)
        tau_as = t1t1 - c("ib,ja->ijab", t1, t1)
        wmnij = (self.blk("oooo")
                 + c("mnie,je->mnij", self.blk("ooov"), t1)
                 - c("mnje,ie->mnij", self.blk("ooov"), t1)
                 - 0.5 * c("ijef,mnfe->ijmn", oovv, tau_as))
        wmbij = (self.blk("ovoo").copy()
                 - c("nb,mnij->mbij", t1, wmnij)
                 + 0.5 * c("mbef,ijef->mbij", self.blk("ovvv"), tau_as))
        zmejb = self.blk("ovvo").copy()
        zmejb_t1 = (c("mbei,je->mbij", self.blk("ovvo"), t1)
                    - c("mbej,ie->mbij", zmejb, t1))
        wmbij = wmbij - zmejb_t1
        rhs = rhs - 0.5 * c("iemn,mnae->ia", wmbij, lam2)
        if self.vchunk is None:
            vvvv_t1 = c("abef,if->abei", self.blk("vvvv"), t1)
        else:
            vvvv_t1 = np.empty((self.nvir, self.nvir, self.nvir, self.nocc), dtype=np.result_type(t1, t2, F))
            for A in self._chunks():
                vvvv_t1[A] = c("abef,if->abei", self.blk("vvvv")[A], t1)
        wabei = self.blk("vvvo").copy() + vvvv_t1
        zmbej = (self.blk("ovvo").copy()
                 + c("mbef,jf->mbej", self.blk("ovvv"), t1)
                 - 0.5 * c("mnej,nb->mbej", self.blk("oovo"), t1)
                 - 0.5 * c("mnef,jnfb->mbej", oovv, t1t1))
        wabei = wabei + c("maei,mb->abei", zmbej, t1)
        wabei = wabei - c("mbei,ma->abei", zmbej, t1)
        return rhs + 0.5 * c("efam,imef->ia", wabei, lam2)

    def L2eq_rhs_CC2(self, t1, t2, lam1, lam2, F):
        """CC2_Helper.py:128-175 (positional einsum outputs kept as written)"""
        c, o, v = self.contract, self.occ, self.vir
        rhs = self.blk("oovv").copy()
        fvv = F[v, v].copy()
        rhs = rhs + c("ijab,eb->ijab", lam2, fvv)
        rhs = rhs - c("ijab,eb->ijba", lam2, fvv)
        foo = F[o, o].copy()
        rhs = rhs - c("imab,jm->ijab", lam2, foo)
        rhs = rhs + c("imab,jm->jiab", lam2, foo)
        wmina = self.LWmina(t1)
        rhs = rhs - c("ijmb,ma->ijab", wmina, lam1)
        rhs = rhs + c("ijmb,ma->ijba", wmina, lam1)
        weifa = self.LWfiea(t1)
        rhs = rhs + c("ejab,ie->ijab", weifa, lam1)
        rhs = rhs - c("ejab,ie->jiab", weifa, lam1)
        fme = self.Fme(t1, t2, F)
        rhs = rhs + c("ia,jb->ijab", lam1, fme)
        rhs = rhs - c("ia,jb->ijba", lam1, fme)
        rhs = rhs - c("ia,jb->jiab", lam1, fme)
        rhs = rhs + c("ia,jb->jiba", lam1, fme)
        return rhs

    # ------------------------------------------------------ density / dipole
    def Dij(self, t1, t2, lam1, lam2):
        """CCSD_Helper.py:984-988"""
        c = self.contract
        return -(c("je,ie->ij", lam1, t1) + 0.5 * c("jmea,imea->ij", lam2, t2))

    def Dab(self, t1, t2, lam1, lam2):
        """CCSD_Helper.py:991-995"""
        c = self.contract
        return c("nb,na->ab", lam1, t1) + 0.5 * c("mneb,mnea->ab", lam2, t2)

    def Dai(self, t1, t2, lam1, lam2):
        """CCSD_Helper.py:998-1006"""
        c = self.contract
        return (c("ia->ai", t1)
                + c("me,miea->ai", lam1, t2)
                - c("me,ma,ie->ai", lam1, t1, t1)
                - 0.5 * c("mnef,mnaf,ie->ai", lam2, t2, t1)
                - 0.5 * c("mnef,inef,ma->ai", lam2, t2, t1))

    def density(self, t1, t2, lam1, lam2):
        """Correlated 1-RDM [[Dij, lam1], [Dai, Dab]] (CCSD_Helper.py:1068-1077)."""
        top = np.hstack((self.Dij(t1, t2, lam1, lam2), lam1))
        bottom = np.hstack((self.Dai(t1, t2, lam1, lam2), self.Dab(t1, t2, lam1, lam2)))
        return np.vstack((top, bottom))

    def dipole_moment(self, t1, t2, lam1, lam2, mu):
        """CCSD_Helper.py:1066-1092: sum_pq mu_pq * D_pq for x, y, z.

        The reference also builds a Hartree-Fock density (:1080-1081) that does
        not enter the result; it is not computed here.  `mu` is (3, n, n).
        """
        corr_p = self.density(t1, t2, lam1, lam2)
        return [self.contract("ij,ij->", mu[k], corr_p) for k in range(3)]
