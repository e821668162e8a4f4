"""CCSD residual program in spin-orbital form (input to the spin-integration compiler).

Equations.  With H~ = exp(-T1) H exp(T1) (a two-body operator whose creators
are transformed by L = 1 - t1 and annihilators by R = 1 + t1), the CCSD
residuals that ccq's src/cc/t{1,2}{A,B,C}_update.f90 compute are

  R1 = <S| H~ (1 + T2) |0>_C,   R2 = <D| H~ (1 + T2 + T2^2/2) |0>_C,

i.e. the CCD equations with dressed f~ and v~.  The program below evaluates
them without ever forming the dressed vvvv block: the driver <ab||ij>~ and
the ladder 1/2 v~_abef t_ij^ef are gathered into

  Q[xy,ij] = A_iA_j(v)[xy,ij] - (ttv part) + 1/2 sum_ef v[xy,ef] tau[ij,ef]
  R2 += C_a C_b Q      (tau = t2 + t1 t1 antisymmetrised)

so the only o^2 v^4 contraction is the bare-integral ladder over tau, as in
any factorised CCSD code.  A_i / C_a are the annihilator / creator
transforms: A_i X[..i..] = X[..i..] + sum_e X[..e..] t1[i,e] and
C_a X[a..] = X[a..] - sum_m t1[m,a] X[m..].

Index letters: i..p occupied, a..h and q..z virtual.  Tensor index order
for amplitudes is occupied first (ccq's t_vec memory layout).
"""

from __future__ import annotations

from .spinint import Program, BlockKey, compile_section, Section


def declare(prog: Program) -> None:
    d = prog.declare
    d('F', 2, 1, kind='base', spin_sort_only=True)            # bare Fock f[p,q]
    d('V', 4, 2, [(0, 1), (2, 3)], kind='base', spin_sort_only=True)  # <pq||rs>
    d('T1', 2, 1, kind='amp')                                  # t1[i,a]
    d('T2', 4, 2, [(0, 1), (2, 3)], kind='amp')                # t2[i,j,a,b]
    d('R1', 2, 1, kind='resid')
    d('R2', 4, 2, [(0, 1), (2, 3)], kind='resid')
    # T1-dependent (dressing) intermediates
    d('H1', 2, 1, kind='dressed')        # sum_kf V[p,k,q,f] t1[k,f]
    d('FF', 2, 1, kind='dressed')        # f + H1
    d('Fd', 2, 1, kind='dressed')        # dressed Fock f~
    d('Xa', 2, 1, kind='dressed')        # A_i (f + H1)[p,i]
    d('Y3', 4, 2, [(0, 1)], kind='dressed')   # A on slot 3
    d('Y4', 4, 2, [(0, 1)], kind='dressed')   # A on slot 4
    d('Vd', 4, 2, [(0, 1), (2, 3)], kind='dressed')   # dressed v~ blocks
    d('Yp', 4, 2, [(0, 1), (2, 3)], kind='dressed')   # A_iA_j v without ttv part
    # T2-dependent intermediates
    d('tau', 4, 2, [(0, 1), (2, 3)])
    d('Q', 4, 2, [(0, 1), (2, 3)])
    d('K1', 4, 2, [(2, 3)])
    d('W4', 4, 2, [(0, 1), (2, 3)])      # W_mnij
    d('Wr', 4, 2)                        # ring W_mbej
    d('Xoo', 2, 1)
    d('Xvv', 2, 1)


def dress_statements(prog: Program):
    s = prog.stmt
    return [
        # f~ = C_p A_q (f + H1)
        s("H1[mi] += V[mnif] * T1[nf]"),
        s("H1[me] += V[mnef] * T1[nf]"),
        s("H1[ai] += V[anif] * T1[nf]"),
        s("H1[ae] += V[anef] * T1[nf]"),
        s("FF[mi] += F[mi]"), s("FF[mi] += H1[mi]"),
        s("FF[me] += F[me]"), s("FF[me] += H1[me]"),
        s("FF[ai] += F[ai]"), s("FF[ai] += H1[ai]"),
        s("FF[ae] += F[ae]"), s("FF[ae] += H1[ae]"),
        s("Fd[mi] += FF[mi]"), s("Fd[mi] += FF[me] * T1[ie]"),
        s("Fd[me] += FF[me]"),
        s("Fd[ae] += FF[ae]"), s("Fd[ae] -= T1[ma] * FF[me]"),
        s("Xa[ai] += FF[ai]"), s("Xa[ai] += FF[ae] * T1[ie]"),
        s("Fd[ai] += Xa[ai]"), s("Fd[ai] -= T1[ma] * Fd[mi]"),
        # v~_oooo = A_i A_j v
        s("Y3[mnij] += V[mnij]"), s("Y3[mnij] += V[mnej] * T1[ie]"),
        s("Y3[mnif] += V[mnif]"), s("Y3[mnif] += V[mnef] * T1[ie]"),
        s("Vd[mnij] += Y3[mnij]"), s("Vd[mnij] += Y3[mnif] * T1[jf]"),
        # v~_ovvo = C_b A_j v
        s("Y4[mbej] += V[mbej]"), s("Y4[mbej] += V[mbef] * T1[jf]"),
        s("Y4[mnej] += V[mnej]"), s("Y4[mnej] += V[mnef] * T1[jf]"),
        s("Vd[mbej] += Y4[mbej]"), s("Vd[mbej] -= T1[nb] * Y4[mnej]"),
        # v~_oovo = A_i v ; v~_ovvv = C_a v
        s("Vd[nmei] += V[nmei]"), s("Vd[nmei] += V[nmef] * T1[if]"),
        s("Vd[maef] += V[maef]"), s("Vd[maef] -= T1[na] * V[mnef]"),
        # driver without its ttv part: Yp[xy,ij] = v + sum_e v[xy,ej] t[i,e] + (i<->j)
        s("Yp[abij] += V[abij]"), s("Yp[abij] += V[abej] * T1[ie]"), s("Yp[abij] += V[abif] * T1[jf]"),
        s("Yp[mbij] += V[mbij]"), s("Yp[mbij] += V[mbej] * T1[ie]"), s("Yp[mbij] += V[mbif] * T1[jf]"),
        s("Yp[mnij] += V[mnij]"), s("Yp[mnij] += V[mnej] * T1[ie]"), s("Yp[mnij] += V[mnif] * T1[jf]"),
    ]


def r2_statements(prog: Program):
    s = prog.stmt
    return [
        s("tau[ijab] += T2[ijab]"),
        s("tau[ijab] += T1[ia] * T1[jb]"),
        s("tau[ijab] -= T1[ib] * T1[ja]"),
        # Q = Yp + 1/2 v tau  (vv, ov, oo creator blocks)
        s("Q[abij] += Yp[abij]"), s("Q[abij] += 1/2 * V[abef] * tau[ijef]"),
        s("Q[mbij] += Yp[mbij]"), s("Q[mbij] += 1/2 * V[mbef] * tau[ijef]"),
        s("Q[mnij] += Yp[mnij]"), s("Q[mnij] += 1/2 * V[mnef] * tau[ijef]"),
        # C_a C_b Q
        s("K1[mbij] += Q[mbij]"), s("K1[mbij] -= T1[nb] * Q[mnij]"),
        s("R2[ijab] += Q[abij]"),
        s("R2[ijab] -= T1[nb] * Q[anij]"),
        s("R2[ijab] -= T1[ma] * K1[mbij]"),
        # Fock-like terms
        s("R2[ijab] += Fd[be] * T2[ijae]"), s("R2[ijab] -= Fd[ae] * T2[ijbe]"),
        s("R2[ijab] -= Fd[mj] * T2[imab]"), s("R2[ijab] += Fd[mi] * T2[jmab]"),
        # hole ladder with its quadratic part
        s("W4[mnij] += Vd[mnij]"), s("W4[mnij] += 1/2 * V[mnef] * T2[ijef]"),
        s("R2[ijab] += 1/2 * W4[mnij] * T2[mnab]"),
        # rings with their quadratic part
        s("Wr[mbej] += Vd[mbej]"), s("Wr[mbej] += 1/2 * V[mnef] * T2[jnbf]"),
        s("R2[ijab] += Wr[mbej] * T2[imae]"),
        s("R2[ijab] -= Wr[maej] * T2[imbe]"),
        s("R2[ijab] -= Wr[mbei] * T2[jmae]"),
        s("R2[ijab] += Wr[maei] * T2[jmbe]"),
        # remaining quadratic terms
        s("Xoo[ni] += 1/2 * V[mnef] * T2[imfe]"),
        s("R2[ijab] -= Xoo[ni] * T2[njab]"), s("R2[ijab] += Xoo[nj] * T2[niab]"),
        s("Xvv[fa] += 1/2 * V[mnef] * T2[nmae]"),
        s("R2[ijab] -= Xvv[fa] * T2[ijfb]"), s("R2[ijab] += Xvv[fb] * T2[ijfa]"),
    ]


def r1_statements(prog: Program):
    s = prog.stmt
    return [
        s("R1[ia] += Fd[ai]"),
        s("R1[ia] += Fd[me] * T2[imae]"),
        s("R1[ia] -= 1/2 * Vd[nmei] * T2[mnae]"),
        s("R1[ia] -= 1/2 * Vd[maef] * T2[imef]"),
    ]


A, B = 0, 1
BLK = {
    't1a': BlockKey('R1', 'ov', (A, A)), 't1b': BlockKey('R1', 'ov', (B, B)),
    't2aa': BlockKey('R2', 'oovv', (A, A, A, A)),
    't2ab': BlockKey('R2', 'oovv', (A, B, A, B)),
    't2bb': BlockKey('R2', 'oovv', (B, B, B, B)),
}


def build(rhf: bool):
    """Compile the ccq-order CCSD sweep (src/update_cc.F90:102-207).

    Returns (prog, sections).  Sections: 'dress' (T1-dependent intermediates),
    residual sections named after the amplitude block they update, 'update'
    and 'copy' markers in ccq's order T2A, T2C|copy, T2B, T1A, T1B|copy.
    """
    prog = Program()
    declare(prog)
    r2 = r2_statements(prog)
    r1 = r1_statements(prog)
    res = {}
    order = ['t2aa'] + ([] if rhf else ['t2bb']) + ['t2ab', 't1a'] + ([] if rhf else ['t1b'])
    for name in order:
        stm = r2 if name.startswith('t2') else r1
        res[name] = compile_section(prog, 'resid:' + name, stm, [BLK[name]])
    dress_stm = dress_statements(prog)

    def dressed_needs(names):
        need = set()
        for n in names:
            for op in res[n].ops:
                for k, _ in op.ins:
                    if prog.decls[k.name].kind == 'dressed':
                        need.add(k)
        return sorted(need, key=lambda k: k.tag())

    sections = []
    top = [n for n in order if n != 't1b']
    sections.append(compile_section(prog, 'dress', dress_stm, dressed_needs(top),
                                    live_kinds=('dressed',)))
    for name in order:
        if name == 't1b':
            sections.append(compile_section(prog, 'dress:t1b', dress_stm, dressed_needs(['t1b']),
                                            live_kinds=('dressed',)))
        sections.append(res[name])
        sections.append(Section(name='update:' + name, kind='update', target=BLK[name],
                                extra={'block': name}))
        if rhf and name == 't2aa':
            sections.append(Section(name='copy:t2bb', kind='copy', extra={'dst': 't2bb', 'src': 't2aa'}))
        if rhf and name == 't1a':
            sections.append(Section(name='copy:t1b', kind='copy', extra={'dst': 't1b', 'src': 't1a'}))
    return prog, sections
