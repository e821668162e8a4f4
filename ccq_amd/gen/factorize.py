"""Binary factorisation of generated residual terms into DSL statements.

Every term is H x T_a [x T_b [x T_c]] where H (the T1-dressed Fock or
two-electron operator) is connected to every T and the T's share no index.
The cheapest order of pairwise contractions (FLOPs estimated at given
spin-orbital sizes) is chosen; H is always contracted first with one T and
the partial result with the next.  Partial results become intermediates,
shared between terms when their definitions coincide after canonical
relabelling.  Intermediate labels are ordered creators first (H creator
slots, T virtual indices) then annihilators (H annihilator slots, T
occupied indices), which fixes their spin-conservation split.
"""

from __future__ import annotations

import itertools
from fractions import Fraction

from .wick import OCC_EXT, VIR_EXT, h_spaces

OCC_DUMMY = "mnop"
VIR_DUMMY = "efghqrstu"


def _space(l: str) -> str:
    return 'o' if l in "ijklmnop" else 'v'


class Factorizer:
    def __init__(self, prog, hnames=None, tnames=None, sizes=None, prefix="X"):
        self.prog = prog
        self.hnames = hnames or {'f': 'Fd', 'v': 'Hv'}
        self.tnames = tnames or {'t2': 'T2', 't3': 'T3', 't4': 'T4'}
        self.sizes = sizes or {'o': 10, 'v': 70}
        self.prefix = prefix
        self.inter = {}      # canonical definition -> (name, label order)
        self.count = 0
        self.needed_h = set()   # (kind, spaces) of H blocks used

    # -----------------------------------------------------------------
    def _letters(self, term):
        m = {}
        po, pv = iter(OCC_DUMMY), iter(VIR_DUMMY)
        sp = h_spaces(term, 0)
        for lab, s in zip(term.h[1], sp):
            if lab.startswith('#'):
                m[lab] = next(po) if s == 'o' else next(pv)
        return m

    def _cost(self, labels):
        c = 2.0
        for l in set(labels):
            c *= self.sizes[_space(l)]
        return c

    def statements(self, term, out_name: str, rank: int):
        """DSL statement strings for one term added into out_name[ijk..abc..]."""
        m = self._letters(term)
        L = lambda x: m.get(x, x)
        sp = h_spaces(term, rank)
        self.needed_h.add((term.h[0], sp))
        hname = self.hnames[term.h[0]]
        h_labels = "".join(L(x) for x in term.h[1])
        # roles of H slots: creators are the first half
        nh = len(h_labels)
        h_role = ['c'] * (nh // 2) + ['a'] * (nh // 2)
        facs = []
        for name, labs in term.ts:
            n = len(labs) // 2
            lab = "".join(L(x) for x in labs)
            facs.append((self.tnames[name], lab, ['a'] * n + ['c'] * n))
        out = "".join(OCC_EXT[:rank]) + "".join(VIR_EXT[:rank])
        coef = term.coef
        if not facs:
            return [f"{out_name}[{out}] += {coef} * {hname}[{h_labels}]"]
        # choose the T order that minimises the contraction cost
        best = None
        for order in itertools.permutations(range(len(facs))):
            cur = set(h_labels)
            cost = 0.0
            for k in order:
                lab = facs[k][1]
                cost += self._cost(cur | set(lab))
                cur = cur.symmetric_difference(set(lab))
            if best is None or cost < best[0]:
                best = (cost, order)
        order = best[1]
        stmts = []
        cur_name, cur_lab, cur_role = hname, h_labels, h_role
        for step, k in enumerate(order):
            tname, tlab, trole = facs[k]
            last = step == len(order) - 1
            if last:
                stmts.append(f"{out_name}[{out}] += {coef} * {cur_name}[{cur_lab}] * {tname}[{tlab}]")
                break
            shared = set(cur_lab) & set(tlab)
            open_ = [(l, r) for l, r in zip(cur_lab, cur_role) if l not in shared] + \
                    [(l, r) for l, r in zip(tlab, trole) if l not in shared]
            cre = [l for l, r in open_ if r == 'c']
            ann = [l for l, r in open_ if r == 'a']
            new_lab = "".join(cre + ann)
            name, fresh = self._intermediate(cur_name, cur_lab, tname, tlab, new_lab, len(cre))
            if fresh:
                stmts.append(f"{name}[{new_lab}] += {cur_name}[{cur_lab}] * {tname}[{tlab}]")
            cur_name, cur_lab = name, new_lab
            cur_role = ['c'] * len(cre) + ['a'] * len(ann)
        return stmts

    def _intermediate(self, a, la, b, lb, lout, nbra):
        """Name for the intermediate lout = a[la] * b[lb] (shared when identical)."""
        # canonical relabelling of the definition
        m = {}
        po, pv = iter("ijklmnop"), iter("abcdefghqrstuvwxyz")
        for l in lout + la + lb:
            if l not in m:
                m[l] = next(po) if _space(l) == 'o' else next(pv)
        key = (a, "".join(m[x] for x in la), b, "".join(m[x] for x in lb), "".join(m[x] for x in lout))
        if key in self.inter:
            return self.inter[key], False
        self.count += 1
        name = f"{self.prefix}{self.count}"
        self.prog.declare(name, len(lout), nbra)
        self.inter[key] = name
        return name, True


def dressed_v_statements(prog, blocks):
    """Statements computing fully T1-dressed v~ blocks 'Hv' from bare V.

    v~ = C_p C_q A_r A_s v with A_r X[..r..] = X[..r..] + sum_e X[..e..] t[r,e]
    (occupied annihilator slots) and C_p X[p..] = X[p..] - sum_m t[m,p] X[m..]
    (virtual creator slots), applied slot 3, 4, 1, 2 in turn.
    ``blocks``: iterable of space strings like 'vvvo' (canonical: o before v
    inside each pair).
    """
    for name, groups in (('D3', [(0, 1)]), ('D4', [(0, 1), (2, 3)]), ('D5', [(2, 3)]),
                         ('Hv', [(0, 1), (2, 3)])):
        if name not in prog.decls:
            prog.declare(name, 4, 2, groups, kind='dressed')
    lab = {'o': iter("ijkl"), 'v': iter("abcd")}
    need = {'D3': set(), 'D4': set(), 'D5': set(), 'Hv': set(blocks)}

    def canon(sp, groups):
        sp = list(sp)
        for g in groups:
            vals = sorted(sp[i] for i in g)
            for i, v in zip(g, vals):
                sp[i] = v
        return "".join(sp)

    # dependencies (closure)
    for sp in list(need['Hv']):
        need['D5'].add(sp)
        if sp[1] == 'v':
            need['D5'].add(sp[0] + 'o' + sp[2:])
    for sp in list(need['D5']):
        need['D4'].add(canon(sp, [(0, 1), (2, 3)]))
        if sp[0] == 'v':
            need['D4'].add(canon('o' + sp[1:], [(0, 1), (2, 3)]))
    for sp in list(need['D4']):
        need['D3'].add(canon(sp, [(0, 1)]))
        if sp[3] == 'o':
            need['D3'].add(canon(sp[:3] + 'v', [(0, 1)]))
    stm = []

    def L(sp):
        it = {'o': iter("ijkl"), 'v': iter("abcd")}
        return "".join(next(it[c]) for c in sp)

    for sp in sorted(need['D3']):
        x = L(sp)
        stm.append(f"D3[{x}] += V[{x}]")
        if sp[2] == 'o':
            e = 'e'
            stm.append(f"D3[{x}] += V[{x[:2]}{e}{x[3]}] * T1[{x[2]}{e}]")
    for sp in sorted(need['D4']):
        x = L(sp)
        stm.append(f"D4[{x}] += D3[{x}]")
        if sp[3] == 'o':
            f = 'f'
            stm.append(f"D4[{x}] += D3[{x[:3]}{f}] * T1[{x[3]}{f}]")
    for sp in sorted(need['D5']):
        x = L(sp)
        stm.append(f"D5[{x}] += D4[{x}]")
        if sp[0] == 'v':
            stm.append(f"D5[{x}] -= T1[m{x[0]}] * D4[m{x[1:]}]")
    for sp in sorted(need['Hv']):
        x = L(sp)
        stm.append(f"Hv[{x}] += D5[{x}]")
        if sp[1] == 'v':
            stm.append(f"Hv[{x}] -= T1[n{x[1]}] * D5[{x[0]}n{x[2:]}]")
    del lab
    return [prog.stmt(s) for s in stm]


def canonical_h_block(sp: str) -> str:
    """o before v inside each antisymmetric pair of a v~ block."""
    a = "".join(sorted(sp[:2]))
    b = "".join(sorted(sp[2:]))
    return a + b
