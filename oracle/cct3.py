"""CC(t;3) tail of ccq (oracle; TEST INFRASTRUCTURE ONLY).

After CCSDt converges, ccq (src/calc_driver.f90:60-72) builds the one- and
two-body similarity-transformed Hamiltonian from T1 and T2 only
(hbar2, src/hbar.f90:10-720), solves the left CCSD equations with it
(solve_lcc / update_l2, src/solve_cc.f90:101-173, src/update_lcc.f90) and
adds the triples correction of crcc23 (src/mm_correct.f90:11-343): a sum over
the triples outside the P space of l3 * M3 / D with four denominators A-D.

This module restates those steps in spin-orbital form with numpy:

* ``dress`` -- T1-similarity-transformed integrals (creators of occupied
  orbitals pick up -t1, annihilators of virtual orbitals +t1);
* ``hbar`` -- the one- and two-body parts of exp(-T2) H~ exp(T2);
* ``lambda_residual`` -- the connected left-CCSD equations written with those
  elements (the three-body parts enter through G_ae, G_mi);
* ``l3``, ``m3`` -- <0|(L1 + L2) Hbar|ijk abc> and <ijk abc|Hbar|0> with
  Hbar from T1, T2;
* ``correction`` -- crcc23's four sums with its denominators, including
  cal_d3's d3d1 loop that starts at occ_a + 2 (src/mm_correct.f90:459).

Spin orbitals are ordered as the brute-force oracle's tensors: occupied
alpha, occupied beta, virtual alpha, virtual beta.  Pins are in
tests/test_oracle_cct3.py (brute-force H-bar matrix, left eigenvector, moments).
"""

from __future__ import annotations

import itertools

import numpy as np

from . import bruteforce as bf


def so_integrals(space: bf.DetSpace, h, eri):
    """Spin-orbital h and <pq||rs> in [occ a, occ b, vir a, vir b] order."""
    hso, vso = bf.spin_orbital_integrals(h, eri)
    order = [space.occ_bit(I) for I in range(space.nocc_so)] + \
            [space.vir_bit(A) for A in range(space.nvir_so)]
    o = np.array(order)
    return hso[np.ix_(o, o)], vso[np.ix_(o, o, o, o)]


def fock(hso, vso, no):
    return hso + np.einsum('pkqk->pq', vso[:, :no, :, :no])


def dress(hso, vso, t1):
    """H~ = exp(-T1) H exp(T1): returns (h~, g~)."""
    no, nv = t1.shape
    N = no + nv
    C = np.eye(N)
    C[:no, no:] = -t1          # I+ -> I+ - sum_A t_IA A+
    A = np.eye(N)
    A[no:, :no] = t1.T         # A -> A + sum_I t_IA I
    ht = C.T @ hso @ A
    gt = np.einsum('pqwx,pr,qs,wu,xv->rsuv', vso, C, C, A, A, optimize=True)
    return ht, gt


class Hbar:
    """One- and two-body blocks of exp(-T2) H~ exp(T2) (T1 dressed in)."""

    def __init__(self, ht, gt, t2, no):
        o, v = slice(0, no), slice(no, None)
        f = ht + np.einsum('pkqk->pq', gt[:, o, :, o])
        g = gt
        self.ov = f[o, v].copy()
        self.vo = None   # singles residual, not used by the left equations
        self.oo = f[o, o] + 0.5 * np.einsum('mnef,inef->mi', g[o, o, v, v], t2)
        self.vv = f[v, v] - 0.5 * np.einsum('mnef,mnaf->ae', g[o, o, v, v], t2)
        self.oooo = g[o, o, o, o] + 0.5 * np.einsum('mnef,ijef->mnij', g[o, o, v, v], t2)
        self.vvvv = g[v, v, v, v] + 0.5 * np.einsum('mnef,mnab->abef', g[o, o, v, v], t2)
        self.ovvo = g[o, v, v, o] + np.einsum('mnef,jnbf->mbej', g[o, o, v, v], t2)
        self.ooov = g[o, o, o, v].copy()
        self.vovv = g[v, o, v, v].copy()
        self.oovv = g[o, o, v, v].copy()
        x = np.einsum('mnie,jnbe->mbij', g[o, o, o, v], t2)
        self.ovoo = (g[o, v, o, o] - np.einsum('me,ijbe->mbij', self.ov, t2)
                     + 0.5 * np.einsum('mbef,ijef->mbij', g[o, v, v, v], t2)
                     + x - x.transpose(0, 1, 3, 2))
        y = np.einsum('mbef,miaf->abei', g[o, v, v, v], t2)
        self.vvvo = (g[v, v, v, o] - np.einsum('me,miab->abei', self.ov, t2)
                     + 0.5 * np.einsum('mnei,mnab->abei', g[o, o, v, o], t2)
                     - y + y.transpose(1, 0, 2, 3))


def _pij(x):
    return x - x.transpose(1, 0, 2, 3)


def _pab(x):
    return x - x.transpose(0, 1, 3, 2)


def lambda_residual(H: Hbar, t2, l1, l2):
    """Connected left-CCSD residuals (l1 [i,a], l2 [i,j,a,b])."""
    Gae = -0.5 * np.einsum('mnef,mnaf->ae', t2, l2)
    Gmi = 0.5 * np.einsum('mnef,inef->mi', t2, l2)
    # write the terms explicitly (W_ieam = ovvo[i,e,a,m])
    r1 = H.ov.copy()
    r1 += np.einsum('ie,ea->ia', l1, H.vv)
    r1 -= np.einsum('ma,im->ia', l1, H.oo)
    r1 += np.einsum('me,ieam->ia', l1, H.ovvo)
    r1 += 0.5 * np.einsum('imef,efam->ia', l2, H.vvvo)
    r1 -= 0.5 * np.einsum('mnae,iemn->ia', l2, H.ovoo)
    r1 -= np.einsum('ef,eifa->ia', Gae, H.vovv)
    r1 -= np.einsum('mn,mina->ia', Gmi, H.ooov)
    r2 = H.oovv.copy()
    r2 += _pab(np.einsum('ijae,eb->ijab', l2, H.vv))
    r2 -= _pij(np.einsum('imab,jm->ijab', l2, H.oo))
    r2 += 0.5 * np.einsum('mnab,ijmn->ijab', l2, H.oooo)
    r2 += 0.5 * np.einsum('ijef,efab->ijab', l2, H.vvvv)
    r2 += _pij(np.einsum('ie,ejab->ijab', l1, H.vovv))
    r2 -= _pab(np.einsum('ma,ijmb->ijab', l1, H.ooov))
    r2 += _pij(_pab(np.einsum('imae,jebm->ijab', l2, H.ovvo)))
    r2 += _pij(_pab(np.einsum('ia,jb->ijab', l1, H.ov)))
    r2 += _pab(np.einsum('ijae,be->ijab', H.oovv, Gae))
    r2 -= _pij(np.einsum('imab,mj->ijab', H.oovv, Gmi))
    return r1, r2


def l3(H: Hbar, l1, l2):
    """<0|(L1 + L2) Hbar|ijk abc> as a full antisymmetric [i,j,k,a,b,c] tensor."""
    O, V = (0, 1, 2), (3, 4, 5)
    x = _p1(_p1(np.einsum('ia,jkbc->ijkabc', l1, H.oovv), O), V)
    x += _p3(_p3(np.einsum('ijab,kc->ijkabc', l2, H.ov), O), V)
    x += _p3(_p1(np.einsum('ijae,ekbc->ijkabc', l2, H.vovv), V), O)
    x -= _p1(_p3(np.einsum('imab,jkmc->ijkabc', l2, H.ooov), V), O)
    return x


def _swap(x, p, q):
    perm = list(range(x.ndim))
    perm[p], perm[q] = perm[q], perm[p]
    return x.transpose(perm)


def _p1(x, ax):
    """P(i/jk): 1 - (ij) - (ik) over three axes."""
    a0, a1, a2 = ax
    return x - _swap(x, a0, a1) - _swap(x, a0, a2)


def _p3(x, ax):
    """P(k/ij): 1 - (ki) - (kj) over three axes."""
    a0, a1, a2 = ax
    return x - _swap(x, a2, a0) - _swap(x, a2, a1)


def left_denominators(H: Hbar):
    """Diagonals of the Hbar one-body blocks (occupied, virtual)."""
    return np.diag(H.oo).copy(), np.diag(H.vv).copy()


def left_jacobi(H: Hbar, t2, l1, l2, ecor, shift=0.0):
    """One update_l2 step (src/update_lcc.f90:10-104, left_jacobi_update
    :208-280): all residuals from the same L, then
    L -= LH / (Hbar_aa - Hbar_ii - E_cor + shift).  Returns the residual
    (ccq's lh_vec) of the step."""
    r1, r2 = lambda_residual(H, t2, l1, l2)
    ho, hv = left_denominators(H)
    d1 = hv[None, :] - ho[:, None] - ecor + shift
    d2 = (hv[None, None, :, None] + hv[None, None, None, :] - ho[:, None, None, None]
          - ho[None, :, None, None] - ecor + shift)
    l1 -= r1 / d1
    l2 -= r2 / d2
    return r1, r2


class SpinMaps:
    """Spin-orbital index <-> (spin, orbital number) for the oracle tensors
    (frozen core not supported: froz = 0)."""

    def __init__(self, norb, oa, ob):
        self.n, self.oa, self.ob = norb, oa, ob
        self.va, self.vb = norb - oa, norb - ob

    def occ(self, spin, i):          # 0-based index in the spin's occupied range
        return i if spin == 0 else self.oa + i

    def vir(self, spin, a):
        return a if spin == 0 else self.va + a


def d3_arrays(H: Hbar, vso_oovv, t2, sm: SpinMaps, d3d1_start=None):
    """cal_d3 (src/mm_correct.f90:345-481) in ccq's spin blocks; arrays are
    indexed with 0-based occupied / virtual indices of their spins.  h2(e,a,j,i)
    of the vvoo block is the bare integral (src/hbar.f90:230, add_twobody_hbar).
    d3d1_start overrides the first beta virtual of d3d1's e loop (ccq's loop
    start is orbital occ_a + 2)."""
    oa, ob, va, vb = sm.oa, sm.ob, sm.va, sm.vb
    no = oa + ob

    def V(spins_v, spins_o, e, a, j, i):
        # <e a || j i> with e, a virtual and j, i occupied
        E, A = sm.vir(spins_v[0], e), sm.vir(spins_v[1], a)
        J, I = sm.occ(spins_o[0], j), sm.occ(spins_o[1], i)
        return vso_oovv[J, I, E, A]

    def T(spins_o, spins_v, i, j, a, b):
        I, J = sm.occ(spins_o[0], i), sm.occ(spins_o[1], j)
        A, B = sm.vir(spins_v[0], a), sm.vir(spins_v[1], b)
        return t2[I, J, A, B]

    d3a1 = np.zeros((va, oa, oa)); d3a2 = np.zeros((va, va, oa))
    d3b1 = np.zeros((va, ob, oa)); d3b2 = np.zeros((vb, va, oa))
    d3c1 = np.zeros((vb, ob, oa)); d3c2 = np.zeros((vb, va, ob))
    d3d1 = np.zeros((vb, ob, ob)); d3d2 = np.zeros((vb, vb, ob))
    for i in range(oa):
        for j in range(oa):
            for a in range(va):
                # ccq t2a(e,a,j,i) = t_{ij}^{ae} (ccq's t2a(a,b,i,j) is t_ij^ab)
                d3a1[a, j, i] = sum(V((0, 0), (0, 0), e, a, j, i) * T((0, 0), (0, 0), i, j, a, e)
                                    for e in range(va))
    for i in range(oa):
        for a in range(va):
            for b in range(va):
                d3a2[b, a, i] = -sum(V((0, 0), (0, 0), b, a, m, i) * T((0, 0), (0, 0), i, m, a, b)
                                     for m in range(oa))
    for i in range(oa):
        for j in range(ob):
            for a in range(va):
                d3b1[a, j, i] = sum(V((1, 0), (1, 0), e, a, j, i) * T((0, 1), (0, 1), i, j, a, e)
                                    for e in range(vb))
    for i in range(oa):
        for a in range(va):
            for b in range(vb):
                d3b2[b, a, i] = -sum(V((1, 0), (1, 0), b, a, m, i) * T((0, 1), (0, 1), i, m, a, b)
                                     for m in range(ob))
    for i in range(oa):
        for j in range(ob):
            for b in range(vb):
                d3c1[b, j, i] = sum(V((1, 0), (1, 0), b, e, j, i) * T((0, 1), (0, 1), i, j, e, b)
                                    for e in range(va))
    for j in range(ob):
        for a in range(va):
            for b in range(vb):
                d3c2[b, a, j] = -sum(V((1, 0), (1, 0), b, a, j, m) * T((0, 1), (0, 1), m, j, a, b)
                                     for m in range(oa))
    # the e loop of d3d1 starts at orbital occ_a + 2 (src/mm_correct.f90:459)
    e0 = max(0, (oa + 1) - ob) if d3d1_start is None else d3d1_start
    for i in range(ob):
        for j in range(ob):
            for a in range(vb):
                d3d1[a, j, i] = sum(V((1, 1), (1, 1), e, a, j, i) * T((1, 1), (1, 1), i, j, a, e)
                                    for e in range(e0, vb))
    for i in range(ob):
        for a in range(vb):
            for b in range(vb):
                d3d2[b, a, i] = -sum(V((1, 1), (1, 1), b, a, m, i) * T((1, 1), (1, 1), i, m, a, b)
                                     for m in range(ob))
    return dict(d3a1=d3a1, d3a2=d3a2, d3b1=d3b1, d3b2=d3b2, d3c1=d3c1, d3c2=d3c2,
                d3d1=d3d1, d3d2=d3d2)


def hbar_diag(H: Hbar, sm: SpinMaps):
    """The Hbar diagonal elements crcc23 uses, as functions of spin-block
    indices: h1(spin, 'o'|'v', p) and h2 for the pair kinds (occ/vir, spins)."""
    def h1(spin, kind, p):
        if kind == 'o':
            I = sm.occ(spin, p)
            return H.oo[I, I]
        A = sm.vir(spin, p)
        return H.vv[A, A]

    def h2_vo(sv, a, so, i):        # h2(a,i,a,i), a virtual, i occupied
        A, I = sm.vir(sv, a), sm.occ(so, i)
        return -H.ovvo[I, A, A, I]

    def h2_oo(s1, j, s2, i):        # h2(j,i,j,i)
        J, I = sm.occ(s1, j), sm.occ(s2, i)
        return H.oooo[J, I, J, I]

    def h2_vv(s1, b, s2, a):        # h2(b,a,b,a)
        B, A = sm.vir(s1, b), sm.vir(s2, a)
        return H.vvvv[B, A, B, A]
    return h1, h2_vo, h2_oo, h2_vv


def tuple_denominators(H: Hbar, fso, sm: SpinMaps, occ, vir, spins_o, spins_v, dd_extra=0.0):
    """crcc23's D_A .. D_D of one triple (src/mm_correct.f90:138-152):
    Fock, H-bar one-body, plus the two-body H-bar diagonals, plus the d3
    three-body part (dd_extra)."""
    h1, h2_vo, h2_oo, h2_vv = hbar_diag(H, sm)
    nv = sm.va + sm.vb
    fo = lambda s, i: fso[sm.occ(s, i), sm.occ(s, i)]
    fv = lambda s, a: fso[len(fso) - nv + sm.vir(s, a), len(fso) - nv + sm.vir(s, a)]
    da = sum(fo(s, x) for s, x in zip(spins_o, occ)) - sum(fv(s, x) for s, x in zip(spins_v, vir))
    db = sum(h1(s, 'o', x) for s, x in zip(spins_o, occ)) - \
        sum(h1(s, 'v', x) for s, x in zip(spins_v, vir))
    dc = db
    for sv, a in zip(spins_v, vir):
        for so, i in zip(spins_o, occ):
            dc += h2_vo(sv, a, so, i)
    for x in range(3):
        for y in range(x + 1, 3):
            dc -= h2_oo(spins_o[y], occ[y], spins_o[x], occ[x])
            dc -= h2_vv(spins_v[y], vir[y], spins_v[x], vir[x])
    return np.array([da, db, dc, dc + dd_extra])


def d3_extra(D, case, occ, vir):
    """The d3 part of crcc23's D_D for one triple (src/mm_correct.f90:116-330):
    sums of cal_d3's arrays over the triple's index pairs; ``case`` is the
    number of beta indices (0 aaa, 1 aab, 2 abb, 3 bbb), occ = (i, j, k),
    vir = (a, b, c) in the block's order."""
    i, j, k = occ
    a, b, c = vir
    if case == 0:
        return (D['d3a1'][a, j, i] + D['d3a1'][a, k, i] + D['d3a1'][a, k, j]
                + D['d3a1'][b, j, i] + D['d3a1'][b, k, i] + D['d3a1'][b, k, j]
                + D['d3a1'][c, j, i] + D['d3a1'][c, k, i] + D['d3a1'][c, k, j]
                - D['d3a2'][b, a, i] - D['d3a2'][b, a, j] - D['d3a2'][b, a, k]
                - D['d3a2'][c, a, i] - D['d3a2'][c, a, j] - D['d3a2'][c, a, k]
                - D['d3a2'][c, b, i] - D['d3a2'][c, b, j] - D['d3a2'][c, b, k])
    if case == 1:
        return (D['d3a1'][a, j, i] + D['d3b1'][a, k, i] + D['d3b1'][a, k, j]
                + D['d3a1'][b, j, i] + D['d3b1'][b, k, i] + D['d3b1'][b, k, j]
                + D['d3c1'][c, k, i] + D['d3c1'][c, k, j]
                - D['d3a2'][b, a, i] - D['d3a2'][b, a, j]
                - D['d3b2'][c, a, i] - D['d3b2'][c, a, j] - D['d3c2'][c, a, k]
                - D['d3b2'][c, b, i] - D['d3b2'][c, b, j] - D['d3c2'][c, b, k])
    if case == 2:
        return (D['d3b1'][a, j, i] + D['d3b1'][a, k, i]
                + D['d3c1'][b, j, i] + D['d3c1'][b, k, i] + D['d3d1'][b, k, j]
                + D['d3c1'][c, j, i] + D['d3c1'][c, k, i] + D['d3d1'][c, k, j]
                - D['d3b2'][b, a, i] - D['d3c2'][b, a, j] - D['d3c2'][b, a, k]
                - D['d3b2'][c, a, i] - D['d3c2'][c, a, j] - D['d3c2'][c, a, k]
                - D['d3d2'][c, b, j] - D['d3d2'][c, b, k])
    return (D['d3d1'][a, j, i] + D['d3d1'][a, k, i] + D['d3d1'][a, k, j]
            + D['d3d1'][b, j, i] + D['d3d1'][b, k, i] + D['d3d1'][b, k, j]
            + D['d3d1'][c, j, i] + D['d3d1'][c, k, i] + D['d3d1'][c, k, j]
            - D['d3d2'][b, a, i] - D['d3d2'][b, a, j] - D['d3d2'][b, a, k]
            - D['d3d2'][c, a, i] - D['d3d2'][c, a, j] - D['d3d2'][c, a, k]
            - D['d3d2'][c, b, i] - D['d3d2'][c, b, j] - D['d3d2'][c, b, k])


CASE_SPINS = {0: ((0, 0, 0), (0, 0, 0)), 1: ((0, 0, 1), (0, 0, 1)),
              2: ((0, 1, 1), (0, 1, 1)), 3: ((1, 1, 1), (1, 1, 1))}


def canonical_triples(sm: SpinMaps, case):
    """(occ, vir) tuples of a spin case in crcc23's loop order: increasing
    inside each same-spin run (i<j<k, a<b<c for aaa; i<j, a<b alpha and k, c
    beta for aab; ...)."""
    oa, ob, va, vb = sm.oa, sm.ob, sm.va, sm.vb
    if case == 0:
        for o in itertools.combinations(range(oa), 3):
            for v in itertools.combinations(range(va), 3):
                yield o, v
    elif case == 1:
        for i, j in itertools.combinations(range(oa), 2):
            for k in range(ob):
                for a, b in itertools.combinations(range(va), 2):
                    for c in range(vb):
                        yield (i, j, k), (a, b, c)
    elif case == 2:
        for i in range(oa):
            for j, k in itertools.combinations(range(ob), 2):
                for a in range(va):
                    for b, c in itertools.combinations(range(vb), 2):
                        yield (i, j, k), (a, b, c)
    else:
        for o in itertools.combinations(range(ob), 3):
            for v in itertools.combinations(range(vb), 3):
                yield o, v


def in_p_space(sm: SpinMaps, case, occ, vir, m1, m2):
    """crcc23's skip rules (src/mm_correct.f90:137,191,247,303): a triple of
    the P space (an active occupied and an active virtual) is skipped."""
    oa, ob = sm.oa, sm.ob
    so, sv = CASE_SPINS[case]
    orb_o = [x + 1 for x in occ]
    orb_v = [(oa if s == 0 else ob) + x + 1 for s, x in zip(sv, vir)]
    i, j, k = orb_o
    a, b, c = orb_v
    if case == 0 or case == 3:
        return k > m1 and a <= m2
    if case == 1:
        return (j > m1 or k > m1) and (a <= m2 or c <= m2)
    return (i > m1 or k > m1) and (a <= m2 or b <= m2)


def correction(H: Hbar, l3so, m3so, fso, d3, sm: SpinMaps, m1: int, m2: int, by_case=False):
    """crcc23's four sums (src/mm_correct.f90:116-330), spin case by spin case,
    over the triples outside the P space; returns (e23a, e23b, e23c, e23d), or
    with by_case the 4 x 4 array [spin case aaa/aab/abb/bbb][A..D].
    Orbital numbers are 1-based with no frozen core; m1, m2 are
    sys%act_occ_b and sys%act_unocc_a."""
    cases = np.zeros((4, 4))
    for case in range(4):
        spins_o, spins_v = CASE_SPINS[case]
        for occ, vir in canonical_triples(sm, case):
            if in_p_space(sm, case, occ, vir, m1, m2):
                continue
            I = [sm.occ(s, x) for s, x in zip(spins_o, occ)]
            A = [sm.vir(s, x) for s, x in zip(spins_v, vir)]
            lm = l3so[I[0], I[1], I[2], A[0], A[1], A[2]] * m3so[I[0], I[1], I[2], A[0], A[1], A[2]]
            den = tuple_denominators(H, fso, sm, occ, vir, spins_o, spins_v,
                                     d3_extra(d3, case, occ, vir))
            cases[case] += lm / den
    if by_case:
        return cases
    return tuple(cases.sum(axis=0))


L_BLOCKS = ['t1a', 't1b', 't2aa', 't2ab', 't2bb']


def _lvec(space, l1, l2):
    b1 = bf.so_to_blocks(space, 1, l1)
    b2 = bf.so_to_blocks(space, 2, l2)
    return np.concatenate([b1[0].ravel(), b1[1].ravel(), b2[0].ravel(), b2[1].ravel(),
                           b2[2].ravel()])


def _set_lvec(space, vec, shapes):
    out, off = [], 0
    for shp in shapes:
        sz = int(np.prod(shp))
        out.append(vec[off:off + sz].reshape(shp))
        off += sz
    l1 = bf.blocks_to_so(space, 1, {0: out[0], 1: out[1]})
    l2 = bf.blocks_to_so(space, 2, {0: out[2], 1: out[3], 2: out[4]})
    return l1, l2


def solve_left(space, H: Hbar, t2so, ecor, tol=1e-7, diis_space=5, max_iter=60, shift=0.0):
    """solve_lcc (src/solve_cc.f90:101-173) with jacobi_iter's loop
    (src/solve_cc.f90:175-269): update_l2 step, pseudo-energy
    sqrt(sum lh^2) + E_cor (src/energy.f90:103-136), the DIIS ring over the
    l_vec layout (T1/T2 blocks).  L starts at zero (ccq allocates l_vec
    without initialising it, src/calc_driver.f90:229-233)."""
    from . import solver as S
    no, nv = space.nocc_so, space.nvir_so
    l1 = np.zeros((no, nv))
    l2 = np.zeros((no, no, nv, nv))
    shapes = [bf.block_shape(1, 0, space.oa, space.ob, space.va, space.vb),
              bf.block_shape(1, 1, space.oa, space.ob, space.va, space.vb),
              bf.block_shape(2, 0, space.oa, space.ob, space.va, space.vb),
              bf.block_shape(2, 1, space.oa, space.ob, space.va, space.vb),
              bf.block_shape(2, 2, space.oa, space.ob, space.va, space.vb)]
    n = diis_space
    ring = [None] * (n + 1)
    conv = [1.0, 1.0, 1.0]
    e_old = 0.0
    hist = []
    converged = False
    for it in range(1, max_iter + 1):
        r1, r2 = left_jacobi(H, t2so, l1, l2, ecor, shift)
        lh = _lvec(space, r1, r2)
        e_new = float(np.sqrt(lh @ lh)) + ecor
        de = e_old - e_new
        conv = [conv[1], conv[2], de]
        e_old = e_new
        slot = it % (n + 1)
        vec = _lvec(space, l1, l2)
        ring[slot - 1 if slot else n] = vec.copy()
        if it % (n + 1) == 0:
            l1, l2 = _set_lvec(space, S.diis_extrapolate(ring, n), shapes)
        hist.append((it, e_new, de))
        if all(abs(c) < tol for c in conv):
            converged = True
            break
    return l1, l2, converged, hist


def cct3(space, h, eri, t_blocks, act, tol=1e-7, diis_space=5):
    """The full CC(t;3) tail on converged CCSDt amplitudes (froz = 0).
    Returns dict(e23=(A, B, C, D), ecor, l1, l2, left_converged)."""
    from . import solver as S
    from . import integrals as I
    no = space.nocc_so
    t1so = bf.blocks_to_so(space, 1, {0: t_blocks['t1a'], 1: t_blocks['t1b']})
    t2so = bf.blocks_to_so(space, 2, {0: t_blocks['t2aa'], 1: t_blocks['t2ab'],
                                      2: t_blocks['t2bb']})
    hso, vso = so_integrals(space, h, eri)
    fso = fock(hso, vso, no)
    ht, gt = dress(hso, vso, t1so)
    H = Hbar(ht, gt, t2so, no)
    e2 = eri.transpose(0, 2, 1, 3)
    fa, fb = I.gen_fock(h, e2, space.oa, space.ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    ecor = S.energy(fa, fb, vaa, vab, vbb, space.oa, space.ob, t_blocks)
    l1, l2, ok, hist = solve_left(space, H, t2so, ecor, tol=tol, diis_space=diis_space)
    l3so = l3(H, l1, l2)
    R, _ = bf.residuals(space, bf.hamiltonian(space, h, eri), {1: t1so, 2: t2so}, [3])
    m3so = R[3]
    sm = SpinMaps(space.n, space.oa, space.ob)
    d3 = d3_arrays(H, vso[:no, :no, no:, no:], t2so, sm)
    e23 = correction(H, l3so, m3so, fso, d3, sm, act.m1, act.m2)
    return dict(e23=e23, ecor=ecor, l1=l1, l2=l2, left_converged=ok, left_history=hist,
                hbar=H, l3=l3so, m3=m3so)
