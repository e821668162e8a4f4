"""Wick-theorem generator of coupled-cluster residual terms (spin-orbital).

Produces the connected terms of  R_mu = <Phi_mu| (H T_n1 T_n2 ...)_C |Phi>
for a general normal-ordered two-body H (one-body part f[p,q] {p+ q}, two-body
part 1/4 v[p,q,r,s] {p+ q+ s r}, every orbital-space block present -- the
T1-similarity-transformed Hamiltonian) and cluster operators T2, T3, T4
(t[i..,a..] with (1/n!)^2 normalisation).  With T1 absorbed in H, these are
all the terms of the CCSDT / CCSDTQ doubles, triples and quadruples residuals.

Method: enumerate every full contraction of
    <0| i1+..ir+ ar..a1 | H | T ... |0>
with Fermi-vacuum quasi-particle contractions (hole creator before hole
annihilator, particle annihilator before particle creator), keep those that
connect H to every T, take the sign from the permutation that brings paired
operators together, and merge equivalent terms with a canonical labelling:
dummy labels are named by the H slot they attach to, T factors are sorted by
what they connect to, antisymmetric groups are put in canonical order with
their permutation parity, and the four permutations of H's slot pairs are
tried.  The output is a list of ``Term``s with Fraction coefficients.
"""

from __future__ import annotations

import dataclasses
import itertools
from fractions import Fraction
from math import factorial

OCC_EXT = "ijkl"
VIR_EXT = "abcd"


def _parity(seq) -> int:
    """Parity of the permutation that sorts ``seq`` (distinct values)."""
    seq = list(seq)
    sign = 1
    for i in range(len(seq)):
        for j in range(i + 1, len(seq)):
            if seq[i] > seq[j]:
                sign = -sign
    return sign


@dataclasses.dataclass(frozen=True)
class Op:
    block: int      # 0 projector, 1 H, 2.. T factors
    space: str      # 'o' or 'v'
    dagger: bool
    label: object   # external letter (projector) or (block, slot) id

    @property
    def left(self) -> bool:
        # hole creator or particle annihilator: left member of a contraction
        return (self.space == 'o' and self.dagger) or (self.space == 'v' and not self.dagger)


@dataclasses.dataclass
class Term:
    coef: Fraction
    h: tuple            # ('f'|'v', labels tuple in slot order)
    ts: tuple           # tuple of ('t<n>', labels tuple: occ..., vir...)

    def __str__(self):
        facs = [f"{self.h[0]}[{''.join(self.h[1])}]"] + [f"{n}[{''.join(l)}]" for n, l in self.ts]
        return f"{self.coef} * " + " * ".join(facs)


def _projector(rank):
    ops = [Op(0, 'o', True, OCC_EXT[k]) for k in range(rank)]
    ops += [Op(0, 'v', False, VIR_EXT[k]) for k in reversed(range(rank))]
    return ops


def _h_ops(kind, spaces):
    if kind == 'f':
        p, q = spaces
        return [Op(1, p, True, (1, 0)), Op(1, q, False, (1, 1))]
    p, q, r, s = spaces
    # p+ q+ s r
    return [Op(1, p, True, (1, 0)), Op(1, q, True, (1, 1)), Op(1, s, False, (1, 3)),
            Op(1, r, False, (1, 2))]


def _t_ops(block, n):
    # a1+ .. an+ in .. i1 ; tensor slots 0..n-1 occupied, n..2n-1 virtual
    ops = [Op(block, 'v', True, (block, n + k)) for k in range(n)]
    ops += [Op(block, 'o', False, (block, k)) for k in reversed(range(n))]
    return ops


def _contractions(ops):
    """Yield lists of (left index, right index) full contractions."""
    n = len(ops)
    lefts = [i for i in range(n) if ops[i].left]
    rights = [i for i in range(n) if not ops[i].left]
    if len(lefts) != len(rights):
        return
    used = [False] * n

    def rec(k, acc):
        if k == len(lefts):
            yield list(acc)
            return
        i = lefts[k]
        for j in rights:
            if used[j] or j < i:
                continue
            if ops[j].space != ops[i].space or ops[j].block == ops[i].block:
                continue
            used[j] = True
            acc.append((i, j))
            yield from rec(k + 1, acc)
            acc.pop()
            used[j] = False

    yield from rec(0, [])


def _canonical(h, ts, externals):
    """Canonical key and sign of a term (h: (kind, labels); ts: [(name, labels)]).

    Labels: external letters, or dummy tuples (block, slot) that appear once in
    H and once in a T factor.
    """
    kind, hl = h
    nsl = len(hl)
    perms = [((0, 1), 1)] if nsl == 2 else []
    if nsl == 4:
        for a, sa in (((0, 1), 1), ((1, 0), -1)):
            for b, sb in (((2, 3), 1), ((3, 2), -1)):
                perms.append((a + b, sa * sb))
    else:
        perms = [((0, 1), 1)]
    best = None
    for order, hsign in perms:
        hlab = [hl[o] for o in order]
        # dummy -> H slot position (canonical name)
        dname = {}
        for pos, lab in enumerate(hlab):
            if not isinstance(lab, str):
                dname[lab] = f"#{pos}"
        hkey = tuple(dname.get(l, l) for l in hlab)
        tdesc = []
        sign = hsign
        for name, labs in ts:
            n = len(labs) // 2
            groups = []
            for g in (labs[:n], labs[n:]):
                named = [dname[l] if not isinstance(l, str) else l for l in g]
                # externals (letters) sort before dummies ('#k'); both by name
                key = sorted(range(len(named)), key=lambda i: (named[i].startswith('#'), named[i]))
                sign *= _parity(key)
                groups.append(tuple(named[i] for i in key))
            tdesc.append((name, groups[0], groups[1]))
        tdesc.sort()
        key = (kind, hkey, tuple(tdesc))
        if best is None or key < best[0]:
            best = (key, sign)
    return best


def residual_terms(rank: int, tprods, hparts=('f', 'v')):
    """Unique terms of <rank| (H * prod T)_C |0> summed over the given T products.

    tprods: list of tuples of T ranks, e.g. [(2,), (3,), (2, 2), (2, 3)].
    Returns a list of Terms (externals i j k.. / a b c..).
    """
    acc = {}
    for tp in tprods:
        mult = 1
        for r in set(tp):
            mult *= factorial(tp.count(r))
        tcoef = Fraction(1, mult)
        for r in tp:
            tcoef /= factorial(r) ** 2
        for kind in hparts:
            hcoef = Fraction(1) if kind == 'f' else Fraction(1, 4)
            nslots = 2 if kind == 'f' else 4
            for spaces in itertools.product('ov', repeat=nslots):
                ops = _projector(rank) + _h_ops(kind, spaces)
                blocks = []
                for k, r in enumerate(tp):
                    blocks.append(2 + k)
                    ops += _t_ops(2 + k, r)
                for pairs in _contractions(ops):
                    # connectivity: H paired with every T block
                    linked = set()
                    for i, j in pairs:
                        bi, bj = ops[i].block, ops[j].block
                        if 1 in (bi, bj):
                            linked.add(bi if bj == 1 else bj)
                    if not all(b in linked for b in blocks):
                        continue
                    seq = []
                    for i, j in pairs:
                        seq += [i, j]
                    sign = _parity(seq)
                    # label identification
                    ident = {}
                    for i, j in pairs:
                        li, lj = ops[i].label, ops[j].label
                        if isinstance(li, str):
                            ident[lj] = li
                        elif isinstance(lj, str):
                            ident[li] = lj
                        else:
                            # H <-> T dummy: name it by the H side
                            hside = li if ops[i].block == 1 else lj
                            ident[li] = hside
                            ident[lj] = hside
                    hl = tuple(ident.get((1, s), (1, s)) for s in range(nslots))
                    ts = []
                    for k, r in enumerate(tp):
                        b = 2 + k
                        ts.append((f"t{r}", tuple(ident[(b, s)] for s in range(2 * r))))
                    key, csign = _canonical((kind, hl), ts, None)
                    acc[key] = acc.get(key, 0) + hcoef * tcoef * sign * csign
    terms = []
    for key, c in acc.items():
        if c == 0:
            continue
        kind, hkey, tdesc = key
        terms.append(Term(Fraction(c), (kind, hkey), tuple((n, g0 + g1) for n, g0, g1 in tdesc)))
    terms.sort(key=lambda t: (len(t.ts), str(t)))
    return terms


def h_spaces(term: Term, rank: int):
    """Orbital space of each H slot of a term."""
    out = []
    for lab in term.h[1]:
        if lab.startswith('#'):
            # a dummy: its space is that of the T slot it connects to
            for name, labs in term.ts:
                n = len(labs) // 2
                if lab in labs:
                    out.append('o' if labs.index(lab) < n else 'v')
                    break
        else:
            out.append('o' if lab in OCC_EXT else 'v')
    return ''.join(out)
