"""Spin-orbital CCSDT residuals at production sizes (oracle; TEST INFRASTRUCTURE ONLY).

ccq computes the CCSDT residual blocks R1 (t1A/t1B), R2 (t2A/t2B/t2C) and
R3 (t3A..t3D) in src/cc/t1A_update.f90 ... src/cc/t3D_update.f90 (generated,
spin-integrated, dgemm + direct antisymmetrised loops, e.g.
src/cc/t3A_update.f90:295-324) and updates them in the order of
src/update_cc.F90:102-318.  Mathematically each block is the projection

    R_mu = <Phi_mu| exp(-T) H exp(T) |Phi>,   T = T1 + T2 + T3,

which this module evaluates in spin orbitals with numpy, at sizes the
brute-force determinant oracle (oracle/bruteforce.py) cannot reach
(40 spatial orbitals, 10 occupied spin orbitals):

* T1 is absorbed by similarity-transforming the integrals,
  H~ = exp(-T1) H exp(T1) (``oracle.cct3.dress``); the residuals are then
  the CCDT projections of H~ with a general (non-diagonal) Fock operator;
* every connected contraction that can reach the projection manifold is
  listed by rank counting (one- and two-body H~ pieces of excitation rank
  -2..+2 with T2, T3, T2^2 and T2 T3), and written as a raw product X_t;
* R_k = sum_t c_t A(X_t), with A the signed sum over the permutations of
  the occupied and of the virtual output indices (no 1/n! factor).

The coefficients c_t are the standard CCSDT ones; they are pinned by
tests/test_oracle_ccsdt_so.py against the brute-force oracle at random
T1, T2, T3 (closed and open shell, every residual element to 1e-12).

R3 is evaluated per canonical occupied triple i<j<k: the occupied part of A
uses each term's own symmetry (a term antisymmetric in two of its occupied
output slots needs three slab evaluations, a fully antisymmetric one needs
one), the virtual part is six transposes of a v^3 slab.  At 40 spatial
orbitals / 10 occupied spin orbitals a residual evaluation is a few hundred
GFLOP through BLAS.

Spin-orbital order (shared with oracle/bruteforce.py): occupied alpha,
occupied beta, virtual alpha, virtual beta; blocks convert with
``oracle.bruteforce.blocks_to_so`` / ``so_to_blocks``.
"""

from __future__ import annotations

import itertools

import numpy as np

from . import bruteforce as bf
from . import cct3


class Dims:
    """Spin-block sizes (the part of bruteforce.DetSpace the block maps use)."""

    def __init__(self, norb, oa, ob):
        self.n, self.oa, self.ob = norb, oa, ob
        self.va, self.vb = norb - oa, norb - ob

    @property
    def nocc_so(self):
        return self.oa + self.ob

    @property
    def nvir_so(self):
        return self.va + self.vb


def so_integrals(h, eri, oa, ob):
    """Spin-orbital h and <pq||rs> in [occ a, occ b, vir a, vir b] order."""
    n = h.shape[0]
    hso, vso = bf.spin_orbital_integrals(h, eri)
    order = np.array(list(range(oa)) + [n + p for p in range(ob)] +
                     list(range(oa, n)) + [n + p for p in range(ob, n)])
    return hso[np.ix_(order, order)], vso[np.ix_(order, order, order, order)]


def _anti(x, groups):
    """Signed sum over the permutations of each axis group (no 1/n!)."""
    for g in groups:
        acc = np.zeros_like(x)
        for p in itertools.permutations(range(len(g))):
            axes = list(range(x.ndim))
            for s, d in zip(g, [g[i] for i in p]):
                axes[s] = d
            acc += bf._parity(p) * x.transpose(axes)
        x = acc
    return x


def _anti_vir3(s):
    """Signed sum over the six permutations of a [a, b, c] slab."""
    return (s - s.transpose(1, 0, 2) - s.transpose(2, 1, 0) - s.transpose(0, 2, 1)
            + s.transpose(1, 2, 0) + s.transpose(2, 0, 1))


class Residuals:
    """T1-dressed spin-orbital CCSDT residuals R1 [i,a], R2 [i,j,a,b],
    R3 [i,j,k,a,b,c] (full antisymmetric tensors)."""

    def __init__(self, h, eri, oa, ob):
        self.dims = Dims(h.shape[0], oa, ob)
        self.no = self.dims.nocc_so
        self.hso, self.vso = so_integrals(h, eri, oa, ob)

    # -- pieces ---------------------------------------------------------------
    def _blocks(self, t1):
        no = self.no
        o, v = slice(0, no), slice(no, None)
        ht, g = cct3.dress(self.hso, self.vso, t1)
        f = ht + np.einsum('pkqk->pq', g[:, o, :, o])
        G = {}
        for key in ('oooo', 'vvvv', 'ovvo', 'oovv', 'vvoo', 'ooov', 'vovv', 'vvvo', 'ovoo'):
            G[key] = np.ascontiguousarray(g[tuple(o if c == 'o' else v for c in key)])
        F = dict(oo=f[o, o].copy(), vv=f[v, v].copy(), ov=f[o, v].copy(), vo=f[v, o].copy())
        return F, G

    def r1(self, F, G, t2, t3):
        """src/cc/t1A_update.f90 (whole T1 residual)."""
        r = F['vo'].T.copy()
        r += np.einsum('me,imae->ia', F['ov'], t2)
        r += 0.5 * np.einsum('amef,imef->ia', G['vovv'], t2, optimize=True)
        r -= 0.5 * np.einsum('mnie,mnae->ia', G['ooov'], t2, optimize=True)
        if t3 is not None:
            r += 0.25 * np.tensordot(t3, G['oovv'], axes=([1, 2, 4, 5], [0, 1, 2, 3]))
        return r

    def r2(self, F, G, t2, t3):
        """src/cc/t2{A,B,C}_update.f90 (whole T2 residual)."""
        grp = [[0, 1], [2, 3]]
        x = 0.25 * G['vvoo'].transpose(2, 3, 0, 1)
        x = x + 0.5 * np.einsum('be,ijae->ijab', F['vv'], t2)
        x -= 0.5 * np.einsum('mj,imab->ijab', F['oo'], t2)
        x += 0.125 * np.einsum('mnij,mnab->ijab', G['oooo'], t2, optimize=True)
        x += 0.125 * np.einsum('abef,ijef->ijab', G['vvvv'], t2, optimize=True)
        x += np.einsum('mbej,imae->ijab', G['ovvo'], t2, optimize=True)
        w = np.einsum('mnef,ijef->mnij', G['oovv'], t2, optimize=True)
        x += 0.0625 * np.einsum('mnij,mnab->ijab', w, t2, optimize=True)
        w = np.einsum('mnef,miea->nfia', G['oovv'], t2, optimize=True)
        x += 0.5 * np.einsum('nfia,njfb->ijab', w, t2, optimize=True)
        w = np.einsum('mnef,mnea->fa', G['oovv'], t2, optimize=True)
        x -= 0.25 * np.einsum('fa,ijfb->ijab', w, t2)
        w = np.einsum('mnef,mief->ni', G['oovv'], t2, optimize=True)
        x -= 0.25 * np.einsum('ni,njab->ijab', w, t2)
        if t3 is not None:
            x += 0.25 * np.tensordot(t3, F['ov'], axes=([2, 5], [0, 1]))
            # sum_mef v(a,m,e,f) t3(i,j,m,e,f,b) -> [i,j,b,a]
            y = np.tensordot(t3, G['vovv'], axes=([2, 3, 4], [1, 2, 3]))
            x -= 0.25 * y.transpose(0, 1, 3, 2)
            # sum_mne v(m,n,i,e) t3(m,n,j,a,b,e) -> [i,j,a,b]
            y = np.tensordot(G['ooov'], t3, axes=([0, 1, 3], [0, 1, 5]))
            x += 0.25 * y
        return _anti(x, grp)

    def r3(self, F, G, t2, t3):
        """src/cc/t3{A..D}_update.f90 (whole T3 residual)."""
        no = self.no
        nv = t2.shape[2]
        oovv = G['oovv']
        # intermediates: H~ contracted with T2 (or with T3) first
        A = np.einsum('mnef,mnae->af', oovv, t2, optimize=True)
        B = np.einsum('mnef,mief->ni', oovv, t2, optimize=True)
        C = np.tensordot(t2, oovv, axes=([0, 1], [0, 1]))            # [a,b,e,f]
        D = np.tensordot(oovv, t2, axes=([2, 3], [2, 3]))            # [m,n,i,j]
        E = np.einsum('mnef,miea->nafi', oovv, t2, optimize=True)     # [n,a,f,i]
        H1 = np.einsum('me,jkec->mjkc', F['ov'], t2)
        I2 = np.einsum('mnie,mnab->ieab', G['ooov'], t2, optimize=True)
        J = np.einsum('mnie,mjea->nija', G['ooov'], t2, optimize=True)
        K = np.einsum('amef,ijef->amij', G['vovv'], t2, optimize=True)
        L = np.einsum('amef,mieb->afib', G['vovv'], t2, optimize=True)
        if t3 is not None:
            t3jk = np.ascontiguousarray(t3.transpose(1, 2, 0, 3, 4, 5))   # [j,k,m,a,b,c]
            t3k = np.ascontiguousarray(t3.transpose(2, 0, 1, 3, 4, 5))    # [k,m,n,a,b,c]
            Fm = np.tensordot(oovv, t3, axes=([1, 2, 3], [0, 3, 4]))   # [m,j,k,c]
            Gm = np.tensordot(oovv, t3, axes=([0, 1, 3], [0, 1, 3]))   # [e,k,b,c]
        # combined operators of equal contraction shape
        Fv = F['vv'] / 12.0 + A / 24.0                                  # full
        Vv = (G['vvvv'] / 24.0 + C / 48.0).reshape(nv * nv, nv * nv)     # full
        Fo = -F['oo'] / 12.0 - B / 24.0                                  # pair23
        Vo = G['oooo'] / 24.0 + D / 48.0                                 # pair12
        Wr = G['ovvo'] / 4.0 + E / 4.0                                   # pair23, [m,a,e,i]
        P1 = -0.25 * H1 + (0.125 * Fm if t3 is not None else 0.0)       # pair23, [m,j,k,c]
        M = 0.5 * L.transpose(2, 0, 3, 1) + 0.125 * I2.transpose(0, 2, 3, 1)  # [i,a,b,f]
        W2 = 0.25 * G['ovoo'] - 0.125 * K.transpose(1, 0, 2, 3) \
            - 0.25 * (J - J.transpose(0, 2, 1, 3)).transpose(0, 3, 1, 2)   # pair12, [m,a,i,j]
        vvvo = G['vvvo']

        def x_pair12(i, j, k):
            """terms antisymmetric in the first two occupied output slots"""
            s = -0.25 * np.tensordot(vvvo[:, :, :, k], t2[i, j], axes=([2], [0]))
            s += np.tensordot(W2[:, :, i, j], t2[:, k], axes=([0], [0]))
            if t3 is not None:
                s += np.tensordot(Vo[:, :, i, j], t3k[k], axes=([0, 1], [0, 1]))
                s += 0.125 * np.tensordot(t2[i, j], Gm[:, k], axes=([1], [0]))
            return s

        def x_pair23(i, j, k):
            """terms antisymmetric in the last two occupied output slots"""
            s = np.tensordot(t2[i], P1[:, j, k], axes=([0], [0])) if t3 is not None else \
                np.tensordot(t2[i], -0.25 * H1[:, j, k], axes=([0], [0]))
            s += np.tensordot(M[i], t2[j, k], axes=([2], [0]))
            if t3 is not None:
                s += np.tensordot(Fo[:, i], t3jk[j, k], axes=([0], [0]))
                s += np.tensordot(Wr[:, :, :, i], t3jk[j, k], axes=([0, 2], [0, 1]))
            return s

        def x_full(i, j, k):
            x = t3[i, j, k]
            s = np.tensordot(Fv, x, axes=([1], [0]))
            s += (Vv @ x.reshape(nv * nv, nv)).reshape(nv, nv, nv)
            return s

        R = np.zeros((no, no, no, nv, nv, nv))
        for i, j, k in itertools.combinations(range(no), 3):
            s = 2.0 * (x_pair12(i, j, k) - x_pair12(k, j, i) - x_pair12(i, k, j))
            s += 2.0 * (x_pair23(i, j, k) - x_pair23(j, i, k) - x_pair23(k, j, i))
            if t3 is not None:
                s += 6.0 * x_full(i, j, k)
            r = _anti_vir3(s)
            for p in itertools.permutations(range(3)):
                idx = tuple((i, j, k)[q] for q in p)
                R[idx] = bf._parity(p) * r
        return R

    # -- public ---------------------------------------------------------------
    def residuals(self, t1, t2, t3=None, ranks=(1, 2, 3)):
        F, G = self._blocks(t1)
        out = {}
        if 1 in ranks:
            out[1] = self.r1(F, G, t2, t3)
        if 2 in ranks:
            out[2] = self.r2(F, G, t2, t3)
        if 3 in ranks:
            out[3] = self.r3(F, G, t2, t3)
        return out

    def to_so(self, t):
        """Spin blocks (t1a, ..., t3bbb) -> spin-orbital T1, T2, T3."""
        d = self.dims
        t1 = bf.blocks_to_so(d, 1, {0: t['t1a'], 1: t['t1b']})
        t2 = bf.blocks_to_so(d, 2, {0: t['t2aa'], 1: t['t2ab'], 2: t['t2bb']})
        t3 = None
        if 't3aaa' in t:
            t3 = bf.blocks_to_so(d, 3, {0: t['t3aaa'], 1: t['t3aab'], 2: t['t3abb'],
                                        3: t['t3bbb']})
        return t1, t2, t3

    def block_residuals(self, t):
        """All residual blocks of spin-block amplitudes ``t`` at once."""
        t1, t2, t3 = self.to_so(t)
        R = self.residuals(t1, t2, t3, ranks=(1, 2, 3) if t3 is not None else (1, 2))
        names = {1: ['t1a', 't1b'], 2: ['t2aa', 't2ab', 't2bb'],
                 3: ['t3aaa', 't3aab', 't3abb', 't3bbb']}
        out = {}
        for rank, r in R.items():
            blk = bf.so_to_blocks(self.dims, rank, r)
            for nb, name in enumerate(names[rank]):
                out[name] = blk[nb]
        return out


class SweepResiduals:
    """Residual provider for ``oracle.solver.sweep`` (ccq's in-place block
    order): re-evaluates whenever the amplitudes changed since the last call."""

    def __init__(self, h, eri, oa, ob):
        self.res = Residuals(h, eri, oa, ob)
        self._key = None
        self._R = None
        self.nevals = 0

    def __call__(self, name, t):
        key = tuple(hash(t[k].tobytes()) for k in sorted(t))
        if key != self._key:
            self._R = self.res.block_residuals(t)
            self._key = key
            self.nevals += 1
        return self._R[name]
