"""Spin-integrated CCSD residual blocks in numpy (oracle).

The projective CCSD residuals R = <mu| exp(-T) H exp(T) |0> that ccq's
generated kernels compute (src/cc/t1A_update.f90, t1B, t2A, t2B, t2C) are
restated here in the T1-similarity-transformed ("dressed") form:

  H~ = exp(-T1) H exp(T1)  is again a two-body operator whose orbitals are
  transformed non-unitarily:  creators by L = 1 - tau, annihilators by
  R = 1 + tau, tau[a, i] = t1[i, a].  Then

  R1 = <S| H~ (1 + T2) |0>_C ,   R2 = <D| H~ (1 + T2 + T2^2/2) |0>_C

are the CCD-form equations with the dressed Fock f~ and integrals v~
(all blocks present; no T1 terms remain).  The spin-orbital CCD and singles
equations are spin-integrated by hand into the alpha/beta blocks below.
These are pinned against the brute-force determinant-space residual
(oracle/bruteforce.py) at random amplitudes.

Layout (C order, occupied indices first, alpha before beta -- the memory
layout of ccq's t_vec blocks, src/utils/cc_utils.f90:479-559):
  t1a[i,a], t1b[i,a], t2aa[i,j,a,b], t2ab[i,j,a,b] (i,a alpha; j,b beta),
  t2bb[i,j,a,b].
Integrals: Vaa/Vbb = antisymmetrised <pq||rs>, Vab = <p q~|r s~>, creators
first (p, q), annihilators last (r, s), over the correlated orbitals.
"""

from __future__ import annotations

import dataclasses

import numpy as np

ES = dict(optimize=True)


@dataclasses.dataclass
class Dressed:
    fa: np.ndarray
    fb: np.ndarray
    Vaa: np.ndarray
    Vab: np.ndarray
    Vbb: np.ndarray
    oa: int
    ob: int

    def sl(self, spin: str, spaces: str):
        """Slice a block: spin in {'aa','ab','bb','a','b'}, spaces like 'ovvo'."""
        occ = {'a': self.oa, 'b': self.ob}
        if len(spin) == 1:
            F = self.fa if spin == 'a' else self.fb
            o = occ[spin]
            idx = tuple(slice(0, o) if c == 'o' else slice(o, None) for c in spaces)
            return F[idx]
        V = {'aa': self.Vaa, 'ab': self.Vab, 'bb': self.Vbb}[spin]
        spins = [spin[0], spin[1], spin[0], spin[1]]
        idx = tuple(slice(0, occ[s]) if c == 'o' else slice(occ[s], None)
                    for s, c in zip(spins, spaces))
        return V[idx]


def _transform_index(X: np.ndarray, axis: int, tau: np.ndarray, o: int, kind: str) -> np.ndarray:
    """Apply L = 1 - tau (kind 'L', creator index) or R = 1 + tau (kind 'R').

    tau is (n, n) with only tau[vir, occ] non-zero.  For a creator index p:
    X'[p] = sum_p' L[p, p'] X[p'] -> X'[vir] -= tau[vir, occ] X[occ].
    For an annihilator index r: X'[r] = sum_r' X[r'] R[r', r] ->
    X'[occ] += sum_vir X[vir] tau[vir, occ].
    """
    Xm = np.moveaxis(X, axis, 0)
    Xm = Xm.copy()
    shp = Xm.shape
    flat = Xm.reshape(shp[0], -1)
    tvo = tau[o:, :o]
    if kind == 'L':
        flat[o:] -= tvo @ flat[:o]
    else:
        flat[:o] += tvo.T @ flat[o:]
    return np.moveaxis(flat.reshape(shp), 0, axis)


def dress(h: np.ndarray, e2int: np.ndarray, t1a: np.ndarray, t1b: np.ndarray,
          froz: int, occ_a: int, occ_b: int) -> Dressed:
    """Build f~ and v~ over the correlated orbitals from bare h and <pq|rs>.

    h, e2int span all orbitals (frozen core included: they enter the Fock sum
    like in src/integrals.f90:327-359); t1 spans correlated occ x virtual.
    """
    n = h.shape[0]
    ta = np.zeros((n, n))
    tb = np.zeros((n, n))
    ta[occ_a:, froz:occ_a] = t1a.T
    tb[occ_b:, froz:occ_b] = t1b.T

    def v4(tp, op, tq, oq):
        X = _transform_index(e2int, 0, tp, op, 'L')
        X = _transform_index(X, 1, tq, oq, 'L')
        X = _transform_index(X, 2, tp, op, 'R')
        return _transform_index(X, 3, tq, oq, 'R')

    Gab = v4(ta, occ_a, tb, occ_b)
    same = occ_a == occ_b and np.array_equal(t1a, t1b)
    Gaa = Gab if same else v4(ta, occ_a, ta, occ_a)
    Gbb = Gab if same else v4(tb, occ_b, tb, occ_b)
    Vaa = Gaa - Gaa.transpose(0, 1, 3, 2)
    Vbb = Gbb - Gbb.transpose(0, 1, 3, 2)
    Vab = Gab

    def h1(tp, o):
        X = _transform_index(h, 0, tp, o, 'L')
        return _transform_index(X, 1, tp, o, 'R')

    fa = h1(ta, occ_a) + np.einsum('pkqk->pq', Vaa[:, :occ_a, :, :occ_a]) \
        + np.einsum('pkqk->pq', Vab[:, :occ_b, :, :occ_b])
    fb = h1(tb, occ_b) + np.einsum('pkqk->pq', Vbb[:, :occ_b, :, :occ_b]) \
        + np.einsum('kpkq->pq', Vab[:occ_a, :, :occ_a, :])
    c = slice(froz, None)
    return Dressed(fa=fa[c, c], fb=fb[c, c],
                   Vaa=np.ascontiguousarray(Vaa[c, c, c, c]),
                   Vab=np.ascontiguousarray(Vab[c, c, c, c]),
                   Vbb=np.ascontiguousarray(Vbb[c, c, c, c]),
                   oa=occ_a - froz, ob=occ_b - froz)


def _asym(X, a1, a2):
    return X - np.swapaxes(X, a1, a2)


def r2_same(d: Dressed, s: str, t2ss, t2ab_s, t2oo):
    """Same-spin doubles residual R[i,j,a,b] for spin s ('a' or 'b').

    t2ab_s is the mixed amplitude in 's-first' order: t_{i m~}^{a e~} with
    i, a of spin s.  t2oo is the opposite same-spin block.
    """
    o = 'b' if s == 'a' else 'a'
    ss, oo = s + s, o + o
    # V blocks in 's-first' ordering: Vso[p,q,r,s] = <p_s q_o | r_s s_o>
    Vso = (lambda spaces: d.sl('ab', spaces)) if s == 'a' else \
        (lambda spaces: d.sl('ab', spaces[1] + spaces[0] + spaces[3] + spaces[2]).transpose(1, 0, 3, 2))
    F = lambda sp: d.sl(s, sp)
    V = lambda sp: d.sl(ss, sp)
    W = lambda sp: d.sl(oo, sp)

    R = V('vvoo').transpose(2, 3, 0, 1).copy()
    R += _asym(np.einsum('be,ijae->ijab', F('vv'), t2ss), 2, 3)
    R += _asym(-np.einsum('mj,imab->ijab', F('oo'), t2ss), 0, 1)
    R += 0.5 * np.einsum('abef,ijef->ijab', V('vvvv'), t2ss, **ES)
    R += 0.5 * np.einsum('mnij,mnab->ijab', V('oooo'), t2ss, **ES)
    ring = np.einsum('mbej,imae->ijab', V('ovvo'), t2ss, **ES) \
        + np.einsum('bmje,imae->ijab', Vso('voov'), t2ab_s, **ES)
    R += _asym(_asym(ring, 0, 1), 2, 3)
    Voovv = V('oovv')
    R += 0.25 * np.einsum('mnef,ijef,mnab->ijab', Voovv, t2ss, t2ss, **ES)
    q2 = np.einsum('mnef,imae,jnbf->ijab', Voovv, t2ss, t2ss, **ES) \
        + np.einsum('mnef,imae,jnbf->ijab', W('oovv'), t2ab_s, t2ab_s, **ES) \
        + np.einsum('mnef,imae,jnbf->ijab', Vso('oovv'), t2ss, t2ab_s, **ES) \
        + np.einsum('nmfe,imae,jnbf->ijab', Vso('oovv'), t2ab_s, t2ss, **ES)
    R += _asym(q2, 0, 1)
    q3 = -0.5 * np.einsum('mnef,imfe,njab->ijab', Voovv, t2ss, t2ss, **ES) \
        - np.einsum('nmfe,imfe,njab->ijab', Vso('oovv'), t2ab_s, t2ss, **ES)
    R += _asym(q3, 0, 1)
    q4 = -0.5 * np.einsum('mnef,nmae,ijfb->ijab', Voovv, t2ss, t2ss, **ES) \
        - np.einsum('nmfe,nmae,ijfb->ijab', Vso('oovv'), t2ab_s, t2ss, **ES)
    R += _asym(q4, 2, 3)
    del t2oo
    return R


def r2aa(d: Dressed, t):
    return r2_same(d, 'a', t['t2aa'], t['t2ab'], t['t2bb'])


def r2bb(d: Dressed, t):
    return r2_same(d, 'b', t['t2bb'], t['t2ab'].transpose(1, 0, 3, 2), t['t2aa'])


def r2ab(d: Dressed, t):
    t2aa, t2ab, t2bb = t['t2aa'], t['t2ab'], t['t2bb']
    Vab = lambda sp: d.sl('ab', sp)
    Vaa = lambda sp: d.sl('aa', sp)
    Vbb = lambda sp: d.sl('bb', sp)
    fa = lambda sp: d.sl('a', sp)
    fb = lambda sp: d.sl('b', sp)
    R = Vab('vvoo').transpose(2, 3, 0, 1).copy()
    R += np.einsum('be,ijae->ijab', fb('vv'), t2ab)
    R += np.einsum('ae,ijeb->ijab', fa('vv'), t2ab)
    R -= np.einsum('mj,imab->ijab', fb('oo'), t2ab)
    R -= np.einsum('mi,mjab->ijab', fa('oo'), t2ab)
    R += np.einsum('abef,ijef->ijab', Vab('vvvv'), t2ab, **ES)
    R += np.einsum('mnij,mnab->ijab', Vab('oooo'), t2ab, **ES)
    # rings
    R += np.einsum('mbej,imae->ijab', Vab('ovvo'), t2aa, **ES)
    R += np.einsum('mbej,imae->ijab', Vbb('ovvo'), t2ab, **ES)
    R -= np.einsum('amej,imeb->ijab', Vab('vovo'), t2ab, **ES)
    R -= np.einsum('mbie,mjae->ijab', Vab('ovov'), t2ab, **ES)
    R += np.einsum('maei,mjeb->ijab', Vaa('ovvo'), t2ab, **ES)
    R += np.einsum('amie,jmbe->ijab', Vab('voov'), t2bb, **ES)
    # quadratic
    Vab_oovv, Vaa_oovv, Vbb_oovv = Vab('oovv'), Vaa('oovv'), Vbb('oovv')
    R += np.einsum('mnef,ijef,mnab->ijab', Vab_oovv, t2ab, t2ab, **ES)
    R += np.einsum('mnef,imae,njfb->ijab', Vaa_oovv, t2aa, t2ab, **ES)
    R += np.einsum('mnef,imae,jnbf->ijab', Vab_oovv, t2aa, t2bb, **ES)
    R += np.einsum('nmfe,imae,njfb->ijab', Vab_oovv, t2ab, t2ab, **ES)
    R += np.einsum('mnef,imae,jnbf->ijab', Vbb_oovv, t2ab, t2bb, **ES)
    R += np.einsum('mnfe,mjae,infb->ijab', Vab_oovv, t2ab, t2ab, **ES)
    R -= 0.5 * np.einsum('mnef,imfe,njab->ijab', Vaa_oovv, t2aa, t2ab, **ES)
    R -= np.einsum('nmfe,imfe,njab->ijab', Vab_oovv, t2ab, t2ab, **ES)
    R -= 0.5 * np.einsum('mnef,jmfe,inab->ijab', Vbb_oovv, t2bb, t2ab, **ES)
    R -= np.einsum('mnfe,mjfe,inab->ijab', Vab_oovv, t2ab, t2ab, **ES)
    R -= 0.5 * np.einsum('mnef,nmae,ijfb->ijab', Vaa_oovv, t2aa, t2ab, **ES)
    R -= np.einsum('nmfe,nmae,ijfb->ijab', Vab_oovv, t2ab, t2ab, **ES)
    R -= 0.5 * np.einsum('mnef,nmbe,ijaf->ijab', Vbb_oovv, t2bb, t2ab, **ES)
    R -= np.einsum('nmef,nmeb,ijaf->ijab', Vab_oovv, t2ab, t2ab, **ES)
    return R


def r1_same(d: Dressed, s: str, t2ss, t2ab_s):
    o = 'b' if s == 'a' else 'a'
    Vso = (lambda spaces: d.sl('ab', spaces)) if s == 'a' else \
        (lambda spaces: d.sl('ab', spaces[1] + spaces[0] + spaces[3] + spaces[2]).transpose(1, 0, 3, 2))
    F = lambda sp: d.sl(s, sp)
    G = lambda sp: d.sl(o, sp)
    V = lambda sp: d.sl(s + s, sp)
    R = F('vo').T.copy()
    R += np.einsum('me,imae->ia', F('ov'), t2ss)
    R += np.einsum('me,imae->ia', G('ov'), t2ab_s)
    R -= 0.5 * np.einsum('nmei,mnae->ia', V('oovo'), t2ss, **ES)
    R -= np.einsum('mnie,mnae->ia', Vso('ooov'), t2ab_s, **ES)
    R -= 0.5 * np.einsum('maef,imef->ia', V('ovvv'), t2ss, **ES)
    R += np.einsum('amef,imef->ia', Vso('vovv'), t2ab_s, **ES)
    return R


def r1a(d: Dressed, t):
    return r1_same(d, 'a', t['t2aa'], t['t2ab'])


def r1b(d: Dressed, t):
    return r1_same(d, 'b', t['t2bb'], t['t2ab'].transpose(1, 0, 3, 2))
