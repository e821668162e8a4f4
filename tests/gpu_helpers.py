"""Shared helpers for the GPU parity tests (oracle on the CPU vs HIP engine)."""

import numpy as np

from oracle import integrals as I
from oracle import solver as osolver
from ccq_amd import engine as E
from ccq_amd import synth


def _groups(shape_name):
    """Antisymmetric axis groups of a block: same-spin runs of occupied and of
    virtual indices (t3aab: occupied (0, 1), virtual (3, 4))."""
    rank, nb = osolver.BLOCKS[shape_name]
    na = rank - nb
    return [list(range(0, na)), list(range(na, rank)),
            list(range(rank, rank + na)), list(range(rank + na, 2 * rank))]


def antisymmetrize_block(x, shape_name):
    """Project onto the antisymmetric subspace of the block's groups."""
    import itertools
    rank = osolver.BLOCKS[shape_name][0]
    for g in _groups(shape_name):
        if len(g) < 2:
            continue
        acc = np.zeros_like(x)
        perms = list(itertools.permutations(range(len(g))))
        for p in perms:
            axes = list(range(2 * rank))
            for src, dst in zip(g, [g[i] for i in p]):
                axes[src] = dst
            acc += osolver.bf._parity(p) * x.transpose(axes)
        x = acc / len(perms)
    return x


def canonical_mask(shape_name, shape):
    """True on tuples strictly increasing inside every antisymmetric group."""
    idx = np.indices(shape)
    mask = np.ones(shape, dtype=bool)
    for g in _groups(shape_name):
        for a, b in zip(g, g[1:]):
            mask &= idx[a] < idx[b]
    return mask


def antisym_random(rng, shape_name, oa, ob, va, vb, scale=0.05):
    x = scale * rng.standard_normal(osolver.block_shape(shape_name, oa, ob, va, vb))
    return antisymmetrize_block(x, shape_name)


class Case:
    """A synthetic system prepared for both the oracle and the engine."""

    def __init__(self, norb, nocc, nL, seed, occ_b=None, froz=0):
        s = synth.make_system(norb, nocc, nL, seed)
        self.sys = s
        self.n, self.oa = norb, nocc
        self.ob = nocc if occ_b is None else occ_b
        self.froz = froz
        e2 = s.e2int()
        self.h, self.e2 = s.h, e2
        self.fa, self.fb = I.gen_fock(s.h, e2, self.oa, self.ob)
        self.vaa, self.vab, self.vbb = I.antisymmetrize(e2, froz)
        self.ca, self.cb = self.oa - froz, self.ob - froz
        self.va, self.vb = norb - self.oa, norb - self.ob

    def engine(self, rhf=False, shift=0.0, diis=5):
        eng = E.Engine(self.froz, self.oa, self.ob, self.n, "CCSD", rhf=rhf, shift=shift,
                       diis_space=diis)
        eng.set_ints(self.fa, self.fb, self.vaa, self.vab, self.vbb)
        return eng

    def provider(self):
        return osolver.CCSDResiduals(self.h, self.e2, self.froz, self.oa, self.ob)

    def zero_t(self):
        return osolver.zero_amplitudes('CCSD', self.ca, self.cb, self.va, self.vb)

    def random_t(self, seed, scale=0.05):
        rng = np.random.default_rng(seed)
        t = self.zero_t()
        for k in ['t1a', 't1b', 't2aa', 't2ab', 't2bb']:
            t[k] = antisym_random(rng, k, self.ca, self.cb, self.va, self.vb, scale)
        return t

    def tvec(self, t):
        return E.join_tvec(t, self.ca, self.cb, self.va, self.vb)

    def blocks(self, vec):
        return E.split_tvec(vec, self.ca, self.cb, self.va, self.vb)

    def oracle_energy(self, t):
        fa = self.fa[self.froz:, self.froz:]
        fb = self.fb[self.froz:, self.froz:]
        return osolver.energy(fa, fb, self.vaa, self.vab, self.vbb, self.ca, self.cb, t)


def ccsdtq_sweep_case():
    """Inputs of the stored CCSDTQ sweep fixture (golden/make_ccsdtq_sweep.py):
    8 orbitals, 4 alpha / 3 beta occupied, random antisymmetric T1-T4."""
    n, oa, ob = 8, 4, 3
    s = synth.make_system(n, oa, 20, 9)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    va, vb = n - oa, n - ob
    rng = np.random.default_rng(2)
    t = osolver.zero_amplitudes('CCSDTQ', oa, ob, va, vb)
    for k in t:
        t[k] = antisym_random(rng, k, oa, ob, va, vb, 0.03)
    return dict(n=n, oa=oa, ob=ob, va=va, vb=vb, h=s.h, eri=s.eri, fa=fa, fb=fb, vaa=vaa,
                vab=vab, vbb=vbb, t=t, shift=0.05)
