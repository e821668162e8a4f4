"""CC(t;3) tail programs in spin-orbital form: H-bar, left CCSD, L3 and M3.

ccq runs, after CCSDt (src/calc_driver.f90:60-72):

* hbar2 (src/hbar.f90:10-720): one- and two-body H-bar = (H exp(T1+T2))_C
  from T1 and T2 only;
* solve_lcc / update_l2 (src/solve_cc.f90:101-173, src/update_lcc.f90): the
  left CCSD equations with that H-bar, Jacobi steps on all L blocks at once,
  denominators Hbar_aa - Hbar_ii - E_cor + shift;
* crcc23 (src/mm_correct.f90:11-343): for every triples spin case, l3 =
  <0|(L1 + L2) Hbar|ijk abc> and the CCSD-type moment M3 = <ijk abc|Hbar|0>
  (T1, T2 only), summed over the triples outside the P space with four
  denominators.

Here H-bar is built from the T1-dressed integrals (f~ = 'Fd', v~ = 'Hv') as
the one- and two-body part of exp(-T2) H~ exp(T2):

  HB1_me = f~_me
  HB1_mi = f~_mi + 1/2 v~_mnef t_inef          HB1_ae = f~_ae - 1/2 v~_mnef t_mnaf
  HB_mnij = v~_mnij + 1/2 v~_mnef t_ijef       HB_abef = v~_abef + 1/2 v~_mnef t_mnab
  HB_mbej = v~_mbej + v~_mnef t_jnbf
  HB_mnie = v~_mnie   HB_amef = v~_amef   HB_mnef = v~_mnef
  HB_mbij = v~_mbij - HB1_me t_ijbe + 1/2 v~_mbef t_ijef + P(ij) v~_mnie t_jnbe
  HB_abei = v~_abei - HB1_me t_miab + 1/2 v~_mnei t_mnab - P(ab) v~_mbef t_miaf

The left residuals are the connected left-CCSD equations in these elements
(G_ae = -1/2 t_mnef l_mnaf, G_mi = 1/2 t_mnef l_inef carry the three-body
part); antisymmetrisers P(ab), P(ij), P(i/jk), ... are written out term by
term, so every spin block (mixed ones included) receives all of them.
"""

from __future__ import annotations

from . import ccsd_program as cp
from . import factorize as fz
from . import wick
from .spinint import BlockKey, Program, Section, compile_section

A, B = 0, 1

# H-bar blocks (canonical space strings) the left equations and L3 read
HB_BLOCKS = ['oooo', 'ooov', 'oovv', 'ovov', 'ovoo', 'ovvv', 'vvov', 'vvvv']
L_NAMES = {(1, 0): 'l1a', (1, 1): 'l1b', (2, 0): 'l2aa', (2, 1): 'l2ab', (2, 2): 'l2bb'}
T3_NAMES = {0: 't3aaa', 1: 't3aab', 2: 't3abb', 3: 't3bbb'}


def declare(prog: Program) -> None:
    cp.declare(prog)
    d = prog.declare
    d('T3', 6, 3, [(0, 1, 2), (3, 4, 5)], kind='amp')
    for name, groups in (('D3', [(0, 1)]), ('D4', [(0, 1), (2, 3)]), ('D5', [(2, 3)]),
                         ('Hv', [(0, 1), (2, 3)])):
        d(name, 4, 2, groups, kind='dressed')
    d('HB1', 2, 1, kind='hbar')
    d('HB', 4, 2, [(0, 1), (2, 3)], kind='hbar')
    d('L1', 2, 1, kind='amp')
    d('L2', 4, 2, [(0, 1), (2, 3)], kind='amp')
    d('LR1', 2, 1, kind='resid')
    d('LR2', 4, 2, [(0, 1), (2, 3)], kind='resid')
    d('Gv', 2, 1)
    d('Go', 2, 1)
    d('L3', 6, 3, [(0, 1, 2), (3, 4, 5)], kind='resid')
    d('M3', 6, 3, [(0, 1, 2), (3, 4, 5)], kind='resid')


def hbar_statements(prog: Program):
    s = prog.stmt
    return [
        s("HB1[me] += Fd[me]"),
        s("HB1[mi] += Fd[mi]"), s("HB1[mi] += 1/2 * Hv[mnef] * T2[inef]"),
        s("HB1[ae] += Fd[ae]"), s("HB1[ae] -= 1/2 * Hv[mnef] * T2[mnaf]"),
        s("HB[mnij] += Hv[mnij]"), s("HB[mnij] += 1/2 * Hv[mnef] * T2[ijef]"),
        s("HB[abef] += Hv[abef]"), s("HB[abef] += 1/2 * Hv[mnef] * T2[mnab]"),
        s("HB[mbej] += Hv[mbej]"), s("HB[mbej] += Hv[mnef] * T2[jnbf]"),
        s("HB[mnie] += Hv[mnie]"),
        s("HB[amef] += Hv[amef]"),
        s("HB[mnef] += Hv[mnef]"),
        s("HB[mbij] += Hv[mbij]"), s("HB[mbij] -= HB1[me] * T2[ijbe]"),
        s("HB[mbij] += 1/2 * Hv[mbef] * T2[ijef]"),
        s("HB[mbij] += Hv[mnie] * T2[jnbe]"), s("HB[mbij] -= Hv[mnje] * T2[inbe]"),
        s("HB[abei] += Hv[abei]"), s("HB[abei] -= HB1[me] * T2[miab]"),
        s("HB[abei] += 1/2 * Hv[mnei] * T2[mnab]"),
        s("HB[abei] -= Hv[mbef] * T2[miaf]"), s("HB[abei] += Hv[maef] * T2[mibf]"),
    ]


def _perm_terms(coef, expr, perms):
    """Expand an antisymmetriser: ``perms`` is a list of (sign, {label: label})
    substitutions applied to the factor labels of ``expr`` (output labels
    stay)."""
    out = []
    for sign, sub in perms:
        facs = []
        for fac in expr.split("*"):
            fac = fac.strip()
            name, labs = fac.split("[")
            labs = "".join(sub.get(c, c) for c in labs.rstrip("]"))
            facs.append(f"{name}[{labs}]")
        c = coef * sign
        out.append(("+=" if c > 0 else "-=", abs(c), " * ".join(facs)))
    return out


ID = [(1, {})]
P_AB = [(1, {}), (-1, {'a': 'b', 'b': 'a'})]
P_IJ = [(1, {}), (-1, {'i': 'j', 'j': 'i'})]
P_IJ_AB = [(si * sa, {**mi, **ma}) for si, mi in P_IJ for sa, ma in P_AB]
# P(i/jk) = 1 - (ij) - (ik); P(k/ij) = 1 - (ki) - (kj); likewise for a, b, c
P_I_JK = [(1, {}), (-1, {'i': 'j', 'j': 'i'}), (-1, {'i': 'k', 'k': 'i'})]
P_K_IJ = [(1, {}), (-1, {'k': 'i', 'i': 'k'}), (-1, {'k': 'j', 'j': 'k'})]
P_A_BC = [(1, {}), (-1, {'a': 'b', 'b': 'a'}), (-1, {'a': 'c', 'c': 'a'})]
P_C_AB = [(1, {}), (-1, {'c': 'a', 'a': 'c'}), (-1, {'c': 'b', 'b': 'c'})]


def _prod(p, q):
    return [(s1 * s2, {**m1, **m2}) for s1, m1 in p for s2, m2 in q]


def _emit(prog, out, terms):
    from fractions import Fraction
    res = []
    for op, coef, expr in terms:
        c = Fraction(coef)
        cs = "" if c == 1 else f"{c} * "
        res.append(prog.stmt(f"{out} {op} {cs}{expr}"))
    return res


def left_statements(prog: Program):
    s = prog.stmt
    r1 = "LR1[ia]"
    r2 = "LR2[ijab]"
    st = [
        s("Gv[ae] -= 1/2 * T2[mnef] * L2[mnaf]"),
        s("Go[mi] += 1/2 * T2[mnef] * L2[inef]"),
        s("LR1[ia] += HB1[ia]"),
        s("LR1[ia] += L1[ie] * HB1[ea]"),
        s("LR1[ia] -= L1[ma] * HB1[im]"),
        s("LR1[ia] += L1[me] * HB[ieam]"),
        s("LR1[ia] += 1/2 * L2[imef] * HB[efam]"),
        s("LR1[ia] -= 1/2 * L2[mnae] * HB[iemn]"),
        s("LR1[ia] -= Gv[ef] * HB[eifa]"),
        s("LR1[ia] -= Go[mn] * HB[mina]"),
        s("LR2[ijab] += HB[ijab]"),
        s("LR2[ijab] += 1/2 * L2[mnab] * HB[ijmn]"),
        s("LR2[ijab] += 1/2 * L2[ijef] * HB[efab]"),
    ]
    del r1
    st += _emit(prog, r2, _perm_terms(1, "L2[ijae] * HB1[eb]", P_AB))
    st += _emit(prog, r2, _perm_terms(-1, "L2[imab] * HB1[jm]", P_IJ))
    st += _emit(prog, r2, _perm_terms(1, "L1[ie] * HB[ejab]", P_IJ))
    st += _emit(prog, r2, _perm_terms(-1, "L1[ma] * HB[ijmb]", P_AB))
    st += _emit(prog, r2, _perm_terms(1, "L2[imae] * HB[jebm]", P_IJ_AB))
    st += _emit(prog, r2, _perm_terms(1, "L1[ia] * HB1[jb]", P_IJ_AB))
    st += _emit(prog, r2, _perm_terms(1, "HB[ijae] * Gv[be]", P_AB))
    st += _emit(prog, r2, _perm_terms(-1, "HB[imab] * Go[mj]", P_IJ))
    return st


def l3_statements(prog: Program):
    out = "L3[ijkabc]"
    st = []
    st += _emit(prog, out, _perm_terms(1, "L1[ia] * HB[jkbc]", _prod(P_I_JK, P_A_BC)))
    st += _emit(prog, out, _perm_terms(1, "L2[ijab] * HB1[kc]", _prod(P_K_IJ, P_C_AB)))
    st += _emit(prog, out, _perm_terms(1, "L2[ijae] * HB[ekbc]", _prod(P_K_IJ, P_A_BC)))
    st += _emit(prog, out, _perm_terms(-1, "L2[imab] * HB[jkmc]", _prod(P_I_JK, P_C_AB)))
    return st


# H-bar blocks whose diagonals enter the CC(t;3) denominators:
# h1(p,p) from the one-body blocks, h2(a,i,a,i) from ovov, h2(j,i,j,i) from
# oooo, h2(b,a,b,a) from vvvv, for every spin combination
DIAG_BLOCKS = [BlockKey('HB1', sp, (s, s)) for sp in ('oo', 'vv') for s in (A, B)] + \
    [BlockKey('HB', 'ovov', (so, sv, so, sv)) for so in (A, B) for sv in (A, B)] + \
    [BlockKey('HB', 'oooo', sp) for sp in ((A, A, A, A), (A, B, A, B), (B, B, B, B))] + \
    [BlockKey('HB', 'vvvv', sp) for sp in ((A, A, A, A), (A, B, A, B), (B, B, B, B))]


def key(name, rank, nb):
    na = rank - nb
    spins = (A,) * na + (B,) * nb
    return BlockKey(name, "o" * rank + "v" * rank, spins + spins)


def build():
    """Programs of the CC(t;3) tail: returns (prog, {'hbar': [...], 'left': [...],
    'cct3': [...]}) section lists."""
    prog = Program()
    declare(prog)
    # M3: the triples projection of (H~ exp(T2))_C without T3 (ccq's mm_t3*
    # use T1 and T2 only)
    fac = fz.Factorizer(prog, sizes={'o': 10, 'v': 70}, prefix="M3X_")
    m3_stm = []
    for t in wick.residual_terms(3, [(2,), (2, 2)]):
        m3_stm += fac.statements(t, "M3", 3)
    m3_stm = [prog.stmt(x) for x in m3_stm]
    hv = set(HB_BLOCKS) | {fz.canonical_h_block(sp) for k, sp in fac.needed_h if k == 'v'}
    dress_stm = cp.dress_statements(prog) + fz.dressed_v_statements(prog, sorted(hv))

    def dressed_needs(sections):
        need = set()
        for sec in sections:
            for op in sec.ops:
                for k, _ in op.ins:
                    if prog.decls[k.name].kind == 'dressed':
                        need.add(k)
        return sorted(need, key=lambda k: k.tag())

    left_res = [compile_section(prog, 'lresid:' + L_NAMES[(r, nb)], left_statements(prog),
                                [key('LR%d' % r, r, nb)]) for r, nb in sorted(L_NAMES)]
    l3 = [compile_section(prog, 'l3:' + T3_NAMES[nb], l3_statements(prog), [key('L3', 3, nb)])
          for nb in range(4)]
    m3 = [compile_section(prog, 'm3:' + T3_NAMES[nb], m3_stm, [key('M3', 3, nb)])
          for nb in range(4)]
    # H-bar blocks: exactly those the left and L3 sections read
    hb_targets = {k for sec in left_res + l3 for op in sec.ops for k, _ in op.ins
                  if prog.decls[k.name].kind == 'hbar'}
    # the diagonals crcc23's denominators use (src/mm_correct.f90:116-330)
    hb_targets |= set(DIAG_BLOCKS)
    hb_targets = sorted(hb_targets, key=lambda k: k.tag())
    hbar = compile_section(prog, 'hbar', hbar_statements(prog), hb_targets, fold_target=False,
                           live_kinds=('hbar',))
    written = {op.out for op in hbar.ops}
    missing = [k.tag() for k in hb_targets if k not in written]
    if missing:
        raise RuntimeError("H-bar blocks read but not produced: " + ", ".join(missing))
    dress_h = compile_section(prog, 'dress:hbar', dress_stm, dressed_needs([hbar] + m3),
                              live_kinds=('dressed',))
    out = {'hbar': [dress_h, hbar], 'left': [], 'cct3': []}
    for sec in left_res:
        out['left'].append(sec)
    for r, nb in sorted(L_NAMES):
        out['left'].append(Section(name='lupdate:' + L_NAMES[(r, nb)], kind='update',
                                   target=key('LR%d' % r, r, nb),
                                   extra={'block': L_NAMES[(r, nb)], 'left': True}))
    for nb in range(4):
        out['cct3'] += [l3[nb], m3[nb],
                        Section(name='corr:' + T3_NAMES[nb], kind='cct3', target=key('L3', 3, nb),
                                extra={'block': T3_NAMES[nb], 'm3': key('M3', 3, nb)})]
    return prog, out
