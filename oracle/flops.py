"""FLOP count of one RK4 step of the reference's propagator (test/measurement infrastructure).

The oracle's equations (oracle/cc_equations.py, which restate the reference
term by term) are evaluated on symbolic tensors that carry only their orbital
classes, whether they are complex, and a structural key.  Every contraction is
costed at its cheapest pairwise order with the letters i j k l m n = o and
a b c d e f = v, and 2 / 4 / 8 real flops per multiply-add for
real x real / real x complex / complex x complex.

  F_ref   every contraction of the 4 chains x 4 stages, as written, plus the dipole;
  F_alg   the same with identical contractions (same string, same operand
          keys) counted once per step: stage 1 of all chains sees the same
          state, the non-propagated amplitudes stay at their step-start values,
          and F takes one value per field time (t, t + dt/2, t + dt).
"""
from __future__ import annotations

import itertools

from oracle.cc_equations import CCEquations

_OCC = set("ijklmn")


class Sym:
    __array_priority__ = 100

    def __init__(self, cls, cplx, key):
        self.cls = cls          # e.g. 'oovv'
        self.cplx = cplx
        self.key = key

    # numpy-like surface used by the equations
    def copy(self):
        return self

    def swapaxes(self, a, b):
        c = list(self.cls)
        c[a], c[b] = c[b], c[a]
        return Sym("".join(c), self.cplx, f"swap({self.key},{a},{b})")

    def __getitem__(self, idx):
        sub = "".join("o" if s.start in (0, None) else "v" for s in idx)
        return Sym(sub, self.cplx, f"{self.key}[{sub}]")

    def _bin(self, other, op):
        if isinstance(other, Sym):
            return Sym(self.cls, self.cplx or other.cplx, f"{op}({self.key},{other.key})")
        cplx = self.cplx or isinstance(other, complex)
        return Sym(self.cls, cplx, f"{op}({self.key},{other!r})")

    def __add__(self, o):
        return self._bin(o, "add")

    __radd__ = __add__

    def __sub__(self, o):
        return self._bin(o, "sub")

    def __rsub__(self, o):
        return self._bin(o, "rsub")

    def __mul__(self, o):
        return self._bin(o, "mul")

    __rmul__ = __mul__

    def __neg__(self):
        return Sym(self.cls, self.cplx, f"neg({self.key})")

    @property
    def real(self):
        return Sym(self.cls, False, f"re({self.key})")


class Counter:
    def __init__(self, o, v):
        self.ext = {"o": o, "v": v}
        self.total = 0.0
        self.seen = {}
        self.unique = 0.0

    def size(self, letters, cls_of):
        n = 1
        for ch in letters:
            n *= self.ext[cls_of[ch]]
        return n

    def contract(self, spec, *ops):
        lhs, out = spec.split("->")
        ins = lhs.split(",")
        cls_of = {}
        for s, op in zip(ins, ops):
            for ch, c in zip(s, op.cls):
                cls_of[ch] = c
        for ch in cls_of:  # letters fix the class in every hot-path string
            cls_of[ch] = "o" if ch in _OCC else "v"

        def pair_cost(x, y, xs, ys):
            letters = set(xs) | set(ys)
            mul = 2 * (1 + x.cplx) * (1 + y.cplx) // 1
            mul = 2 if not (x.cplx or y.cplx) else (4 if x.cplx != y.cplx else 8)
            return mul * self.size(letters, cls_of)

        if len(ops) == 1:
            cost = 0.0
        elif len(ops) == 2:
            cost = pair_cost(ops[0], ops[1], ins[0], ins[1])
        else:
            best = None
            for a, b in itertools.combinations(range(3), 2):
                c = 3 - a - b
                mid = "".join(ch for ch in dict.fromkeys(ins[a] + ins[b]) if ch in ins[c] or ch in out)
                midsym = Sym("".join(cls_of[ch] for ch in mid), ops[a].cplx or ops[b].cplx, "")
                cost = pair_cost(ops[a], ops[b], ins[a], ins[b]) + pair_cost(midsym, ops[c], mid, ins[c])
                best = cost if best is None else min(best, cost)
            cost = best
        key = spec + "|" + "|".join(op.key for op in ops)
        self.total += cost
        if key not in self.seen:
            self.seen[key] = cost
            self.unique += cost
        cplx = any(op.cplx for op in ops)
        return Sym("".join(cls_of[ch] for ch in out), cplx, "C(" + key + ")")


class _SymEquations(CCEquations):
    def __init__(self, counter):
        self.contract = counter.contract
        self.occ = slice(0, 1)
        self.vir = slice(1, 2)
        self.vchunk = None  # whole v^4 intermediates, as written

    def blk(self, key):
        return Sym(key, False, "TEI" + key)


def step_flops(o, v, method="CCSD"):
    """(F_ref, F_alg) of one RK4 step of the given method at (o, v)."""
    cnt = Counter(o, v)
    eq = _SymEquations(cnt)
    t1, t2 = Sym("ov", True, "t1"), Sym("oovv", True, "t2")
    l1, l2 = Sym("ov", True, "l1"), Sym("oovv", True, "l2")
    Fs = [Sym("nn", True, f"F@{s}") for s in ("t", "t+dt/2", "t+dt/2", "t+dt")]
    # slicing F uses (slice, slice) with the occ/vir slices of the equations
    if method == "CCSD":
        rhs = (eq.T1eq_rhs, eq.T2eq_rhs, eq.lam_1eq_rhs, eq.lam2eq_rhs)
    else:
        rhs = (eq.T1eq_rhs, eq.T2eq_rhs_CC2, eq.L1eq_rhs_CC2, eq.L2eq_rhs_CC2)
    ys = (t1, t2, l1, l2)
    for chain in range(4):
        k = None
        for s in range(4):
            y = ys[chain] if s == 0 else ys[chain] + (0.5 if s < 3 else 1.0) * k
            args = [t1, t2, l1, l2]
            args[chain] = y
            if chain < 2:
                k = rhs[chain](args[0], args[1], Fs[s])
            else:
                k = rhs[chain](*args, Fs[s])
    for f in (eq.Dij, eq.Dab, eq.Dai):
        f(t1, t2, l1, l2)
    return cnt.total, cnt.unique


def rhs_flops(o, v, name):
    """Flops of one evaluation of the named right-hand side (e.g. 'T2eq_rhs',
    'lam2eq_rhs') at (o, v), each contraction costed as above."""
    cnt = Counter(o, v)
    eq = _SymEquations(cnt)
    t1, t2 = Sym("ov", True, "t1"), Sym("oovv", True, "t2")
    l1, l2 = Sym("ov", True, "l1"), Sym("oovv", True, "l2")
    F = Sym("nn", True, "F")
    if name in ("T1eq_rhs", "T2eq_rhs", "T2eq_rhs_CC2"):
        getattr(eq, name)(t1, t2, F)
    else:
        getattr(eq, name)(t1, t2, l1, l2, F)
    return cnt.total
