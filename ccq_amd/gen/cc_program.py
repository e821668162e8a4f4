"""Full CC programs (CCSD, CCSDT, CCSDTQ) in ccq's update order.

CCSD residual statements are the hand-factorised ones of ``ccsd_program``;
every term involving T3 or T4 -- their contributions to R1 and R2 and the
R3 / R4 residuals themselves -- is produced by the Wick generator
(``wick.py``) with the T1-dressed Hamiltonian (Fock f~ = 'Fd', two-electron
v~ = 'Hv', all blocks) and factorised by ``factorize.py``.

Sweep order and RHF shortcuts follow src/update_cc.F90:102-318:
T2A, T2C|copy, T2B, T1A, T1B|copy, T3A, T3D|copy, T3B, T3C|tran3bto3c,
T4A, T4B, T4C, T4D|tran4bto4d, T4E|copy.
"""

from __future__ import annotations

from . import ccsd_program as cp
from . import factorize as fz
from . import wick
from .spinint import BlockKey, BlockOp, Program, Section, compile_section

A, B = 0, 1

LEVEL = {'CCSD': 2, 'CCSDT': 3, 'CCSDTQ': 4}

# residual T-products per projection rank and method level (T1 absorbed in H~)
TPRODS = {
    3: {1: [(3,)], 2: [(3,)], 3: [(2,), (3,), (2, 2), (2, 3)]},
    4: {1: [(3,)], 2: [(3,), (4,)], 3: [(2,), (3,), (4,), (2, 2), (2, 3)],
        4: [(3,), (4,), (2, 2), (2, 3), (2, 4), (3, 3), (2, 2, 2)]},
}


def amp_key(rank, nb):
    na = rank - nb
    spins = (A,) * na + (B,) * nb
    return BlockKey(f"T{rank}", "o" * rank + "v" * rank, spins + spins)


def res_key(rank, nb):
    na = rank - nb
    spins = (A,) * na + (B,) * nb
    return BlockKey(f"R{rank}", "o" * rank + "v" * rank, spins + spins)


BLOCK_NAMES = {
    (1, 0): 't1a', (1, 1): 't1b',
    (2, 0): 't2aa', (2, 1): 't2ab', (2, 2): 't2bb',
    (3, 0): 't3aaa', (3, 1): 't3aab', (3, 2): 't3abb', (3, 3): 't3bbb',
    (4, 0): 't4aaaa', (4, 1): 't4aaab', (4, 2): 't4aabb', (4, 3): 't4abbb', (4, 4): 't4bbbb',
}
NAME_BLOCK = {v: k for k, v in BLOCK_NAMES.items()}


def _perm_copy_section(name, dst, src, perm_labels):
    """RHF permuted copy between amplitude blocks (tran3bto3c / tran4bto4d)."""
    rd, nbd = NAME_BLOCK[dst]
    rs, nbs = NAME_BLOCK[src]
    out_lab = "ijkl"[:rd] + "abcd"[:rd]
    op = BlockOp(amp_key(rd, nbd), out_lab, 1, ((amp_key(rs, nbs), perm_labels),), assign=True)
    return Section(name=name, kind='compute', ops=[op])


# update_clusters_t3_opt (src/update_cc.F90:325-2191): T3 residuals in three
# groups, each formed from the state at the start of the group, with masked
# updates (block, (vneed, vforbid)): an occupied index and a virtual index of
# a spin in vneed must be active, no virtual of a spin in vforbid may be.
CCSDT_ACT_T12 = ['t2ab', 't2aa', 't2bb', 't1a', 't1b']
CCSDT_ACT_GROUPS = [
    [('t3aab', (1, 0)), ('t3abb', (2, 0))],     # t3BC_update_opt.f: B1, B4, C1, C4
    [('t3aaa', (3, 0)), ('t3aab', (2, 1))],     # t3AB_update_opt.f: A, B2, B3
    [('t3abb', (1, 2)), ('t3bbb', (3, 0))],     # t3CD_update_opt.f: C2, C3, D
]


def build(method: str, rhf: bool):
    active = method == 'CCSDt'
    if active:
        # the active-space path has no closed-shell shortcuts in ccq
        method, rhf = 'CCSDT', False
    lvl = LEVEL[method]
    if lvl == 2:
        return cp.build(rhf)
    prog = Program()
    cp.declare(prog)
    d = prog.declare
    for r in range(3, lvl + 1):
        groups = [tuple(range(r)), tuple(range(r, 2 * r))]
        d(f"T{r}", 2 * r, r, groups, kind='amp')
        d(f"R{r}", 2 * r, r, groups, kind='resid')
    for name, groups in (('D3', [(0, 1)]), ('D4', [(0, 1), (2, 3)]), ('D5', [(2, 3)]),
                         ('Hv', [(0, 1), (2, 3)])):
        d(name, 4, 2, groups, kind='dressed')
    gen = {}
    needed_h = set()
    for rank, tps in TPRODS[lvl].items():
        # one factoriser per residual rank: each residual section defines every
        # intermediate it reads
        fac = fz.Factorizer(prog, sizes={'o': 10, 'v': 70}, prefix=f"X{rank}_")
        terms = wick.residual_terms(rank, tps)
        stm = []
        for t in terms:
            stm += fac.statements(t, f"R{rank}", rank)
        gen[rank] = [prog.stmt(s) for s in stm]
        needed_h |= fac.needed_h
    hv_blocks = sorted({fz.canonical_h_block(sp) for k, sp in needed_h if k == 'v'})
    dress_stm = cp.dress_statements(prog) + fz.dressed_v_statements(prog, hv_blocks)
    r2 = cp.r2_statements(prog) + gen.get(2, [])
    r1 = cp.r1_statements(prog) + gen.get(1, [])
    stm_of = {1: r1, 2: r2}
    for r in range(3, lvl + 1):
        stm_of[r] = gen[r]

    # ccq's sweep order (src/update_cc.F90:102-318)
    order = ['t2aa'] + ([] if rhf else ['t2bb']) + ['t2ab', 't1a'] + ([] if rhf else ['t1b'])
    order += ['t3aaa'] + ([] if rhf else ['t3bbb']) + ['t3aab'] + ([] if rhf else ['t3abb'])
    if lvl >= 4:
        order += ['t4aaaa', 't4aaab', 't4aabb'] + ([] if rhf else ['t4abbb', 't4bbbb'])
    res = {}
    for name in order:
        rank, nb = NAME_BLOCK[name]
        res[name] = compile_section(prog, 'resid:' + name, stm_of[rank], [res_key(rank, nb)],
                                    merge_shared=6)

    def dressed_needs(names):
        need = set()
        for n in names:
            for op in res[n].ops:
                for k, _ in op.ins:
                    if prog.decls[k.name].kind == 'dressed':
                        need.add(k)
        return sorted(need, key=lambda k: k.tag())

    # T1 changes at the T1A and T1B updates: dress for the T2/T1A sections,
    # redress for T1B (uhf), and again for everything after T1 (the T3/T4
    # sections; that state also serves the next sweep's T2 sections).
    pre = [n for n in order if n.startswith('t2') or n == 't1a']
    post = [n for n in order if n.startswith('t3') or n.startswith('t4')]
    if active:
        return prog, _active_sections(prog, res, dress_stm, dressed_needs, pre, post)
    sections = [compile_section(prog, 'dress', dress_stm, dressed_needs(pre),
                                live_kinds=('dressed',))]
    for name in order:
        if name == 't1b':
            sections.append(compile_section(prog, 'dress:t1b', dress_stm, dressed_needs(['t1b']),
                                            live_kinds=('dressed',)))
        if name == post[0]:
            sections.append(compile_section(prog, 'dress:t3', dress_stm, dressed_needs(post),
                                            live_kinds=('dressed',)))
        rank, nb = NAME_BLOCK[name]
        sections.append(res[name])
        sections.append(Section(name='update:' + name, kind='update', target=res_key(rank, nb),
                                extra={'block': name}))
        if rhf:
            if name == 't2aa':
                sections.append(Section(name='copy:t2bb', kind='copy', extra={'dst': 't2bb', 'src': 't2aa'}))
            if name == 't1a':
                sections.append(Section(name='copy:t1b', kind='copy', extra={'dst': 't1b', 'src': 't1a'}))
            if name == 't3aaa':
                sections.append(Section(name='copy:t3bbb', kind='copy', extra={'dst': 't3bbb', 'src': 't3aaa'}))
            if name == 't3aab':
                # src/cc/tran3Bto3C.f90: t3c(c,b,a,k,j,i) = t3b(a,c,b,i,k,j)
                sections.append(_perm_copy_section('copy:t3abb', 't3abb', 't3aab', 'jkibca'))
            if name == 't4aaab':
                pass
            if name == 't4aabb':
                # src/cc/tran4Bto4D.f: T4D from T4B; T4E = T4A (closed shell)
                sections.append(_perm_copy_section('copy:t4abbb', 't4abbb', 't4aaab', 'jklibcda'))
                sections.append(Section(name='copy:t4bbbb', kind='copy',
                                        extra={'dst': 't4bbbb', 'src': 't4aaaa'}))
    return prog, sections


def _active_sections(prog, res, dress_stm, dressed_needs, pre, post):
    """Section list of the CCSDt sweep (T2B, T2A, T2C, T1A, T1B, T3 groups)."""
    def upd(name, mask=None):
        rank, nb = NAME_BLOCK[name]
        extra = {'block': name}
        if mask is not None:
            extra['mask'] = mask
        return Section(name='update:' + name, kind='update', target=res_key(rank, nb), extra=extra)

    sections = [compile_section(prog, 'dress', dress_stm, dressed_needs(pre),
                                live_kinds=('dressed',))]
    for name in CCSDT_ACT_T12:
        if name == 't1b':
            sections.append(compile_section(prog, 'dress:t1b', dress_stm, dressed_needs(['t1b']),
                                            live_kinds=('dressed',)))
        sections += [res[name], upd(name)]
    sections.append(compile_section(prog, 'dress:t3', dress_stm, dressed_needs(post),
                                    live_kinds=('dressed',)))
    for group in CCSDT_ACT_GROUPS:
        sections += [res[name] for name, _ in group]
        sections += [upd(name, mask) for name, mask in group]
    return sections
