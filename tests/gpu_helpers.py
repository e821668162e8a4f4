"""Shared helpers for the GPU parity tests (oracle on the CPU vs HIP engine)."""

import numpy as np

from oracle import integrals as I
from oracle import solver as osolver
from ccq_amd import engine as E
from ccq_amd import synth


def antisym_random(rng, shape_name, oa, ob, va, vb, scale=0.05):
    import itertools
    rank, nb = osolver.BLOCKS[shape_name]
    x = scale * rng.standard_normal(osolver.block_shape(shape_name, oa, ob, va, vb))
    na = rank - nb
    groups = [list(range(0, na)), list(range(na, rank)),
              list(range(rank, rank + na)), list(range(rank + na, 2 * rank))]
    for g in groups:
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
