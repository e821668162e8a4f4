"""Active-space CCSDt (oracle; TEST INFRASTRUCTURE ONLY).

Restates ccq's CCSDt path:

* active orbitals (src/parser.f90:496-499): m1 = max(occ_b - act_occ, froz),
  m2 = min(occ_a + act_unocc, orbs); active occupied are orbitals m1+1..occ
  of either spin, active virtuals occ+1..m2;
* the P space of triples: at least one active occupied and at least one
  active virtual index (act_ind_t = 1, src/parser.f90:380-390;
  src/mm_correct.f90:137);
* the update order of update_clusters_t3_opt (src/update_cc.F90:325-2191):
  T2B, T2A, T2C, T1A, T1B, then three groups whose residuals are formed from
  the state at the start of the group and updated together --
  {T3B1, T3B4, T3C1, T3C4} (t3BC_update_opt.f), {T3A, T3B2, T3B3}
  (t3AB_update_opt.f), {T3C2, T3C3, T3D} (t3CD_update_opt.f);
* the packed t_vec of ten P-space T3 blocks (src/utils/cc_utils.f90:560-631,
  with the index ranges declared in src/cc_opt_ccsdt/t2B0000_update_opt.f:48-57),
  which is also the DIIS vector (conv%vec_size = t_size,
  src/solve_cc.f90:68-77).

Amplitudes are kept in full antisymmetric storage (the C-order blocks of
oracle/solver.py) with the Q-space triples held at zero.
"""

from __future__ import annotations

import dataclasses
import itertools
import math

import numpy as np

from . import solver as S


@dataclasses.dataclass
class ActiveSpace:
    froz: int
    occ_a: int
    occ_b: int
    orbs: int
    act_occ: int
    act_unocc: int

    @property
    def m1(self):
        return max(self.occ_b - self.act_occ, self.froz)

    @property
    def m2(self):
        return min(self.occ_a + self.act_unocc, self.orbs)

    def occ_start(self):
        """First active occupied index (0-based in the correlated range), same for both spins."""
        return self.m1 - self.froz

    def vir_end(self, spin):
        """Number of active virtuals of a spin (they are the lowest ones)."""
        occ = self.occ_b if spin else self.occ_a
        return max(self.m2 - occ, 0)


def _spins(name):
    """Spin (0 alpha, 1 beta) of each occupied / virtual index of a block."""
    rank, nb = S.BLOCKS[name]
    return [0] * (rank - nb) + [1] * nb


def p_mask(name, shape, act: ActiveSpace, vneed=3, vforbid=0):
    """Tuples of a T3 block updated by a masked step.

    Occupied: at least one active index.  Virtual: at least one active index of
    a spin in ``vneed`` (bit 0 alpha, bit 1 beta) and none of a spin in
    ``vforbid``.  vneed=3, vforbid=0 is the P space.
    """
    rank = len(shape) // 2
    spins = _spins(name)
    idx = np.indices(shape)
    occ_act = np.zeros(shape, bool)
    vir_act = [np.zeros(shape, bool), np.zeros(shape, bool)]
    for d in range(rank):
        occ_act |= idx[d] >= act.occ_start()
        s = spins[d]
        vir_act[s] = vir_act[s] | (idx[rank + d] < act.vir_end(s))
    need = np.zeros(shape, bool)
    for s in (0, 1):
        if vneed >> s & 1:
            need |= vir_act[s]
    mask = occ_act & need
    for s in (0, 1):
        if vforbid >> s & 1:
            mask &= ~vir_act[s]
    return mask


# update_clusters_t3_opt: T1/T2 one block at a time, then the T3 groups with
# (block, vneed, vforbid) per masked update
SWEEP = [
    [('t2ab', None)], [('t2aa', None)], [('t2bb', None)], [('t1a', None)], [('t1b', None)],
    # t3BC_update_opt.f: B1 u B4 (an alpha virtual active), C1 u C4 (a beta virtual active)
    [('t3aab', (1, 0)), ('t3abb', (2, 0))],
    # t3AB_update_opt.f: T3A; B2 u B3 (only the beta virtual active)
    [('t3aaa', (3, 0)), ('t3aab', (2, 1))],
    # t3CD_update_opt.f: C2 u C3 (only the alpha virtual active); T3D
    [('t3abb', (1, 2)), ('t3bbb', (3, 0))],
]


def sweep(t, residual, D, act: ActiveSpace):
    """One update_clusters_t3_opt sweep (src/update_cc.F90:325-2191)."""
    for group in SWEEP:
        R = {name: residual(name, t) for name, _ in group}
        for name, m in group:
            if m is None:
                t[name] -= R[name] / D[name]
            else:
                mask = p_mask(name, t[name].shape, act, *m)
                t[name][mask] -= R[name][mask] / D[name][mask]


# ---------------------------------------------------------------------------
# packed t_vec (src/utils/cc_utils.f90:560-631)
#
# Each block lists its six Fortran index ranges (fastest first) as
# (space, spin, part), part in all / act / inact, and the full block it is
# cut from.  Fortran slot s is axis 5 - s of the C-order full block.

PACKED = [
    ('t3aaa', [('v', 0, 'all'), ('v', 0, 'all'), ('v', 0, 'act'),
               ('o', 0, 'all'), ('o', 0, 'all'), ('o', 0, 'act')]),           # t3A
    ('t3aab', [('v', 1, 'all'), ('v', 0, 'all'), ('v', 0, 'act'),
               ('o', 1, 'all'), ('o', 0, 'all'), ('o', 0, 'act')]),           # t3B1
    ('t3aab', [('v', 1, 'act'), ('v', 0, 'inact'), ('v', 0, 'inact'),
               ('o', 1, 'act'), ('o', 0, 'inact'), ('o', 0, 'inact')]),       # t3B2
    ('t3aab', [('v', 1, 'act'), ('v', 0, 'inact'), ('v', 0, 'inact'),
               ('o', 1, 'all'), ('o', 0, 'all'), ('o', 0, 'act')]),           # t3B3
    ('t3aab', [('v', 1, 'all'), ('v', 0, 'all'), ('v', 0, 'act'),
               ('o', 1, 'act'), ('o', 0, 'inact'), ('o', 0, 'inact')]),       # t3B4
    ('t3abb', [('v', 1, 'all'), ('v', 1, 'act'), ('v', 0, 'all'),
               ('o', 1, 'all'), ('o', 1, 'act'), ('o', 0, 'all')]),           # t3C1
    ('t3abb', [('v', 1, 'inact'), ('v', 1, 'inact'), ('v', 0, 'act'),
               ('o', 1, 'inact'), ('o', 1, 'inact'), ('o', 0, 'act')]),       # t3C2
    ('t3abb', [('v', 1, 'inact'), ('v', 1, 'inact'), ('v', 0, 'act'),
               ('o', 1, 'all'), ('o', 1, 'act'), ('o', 0, 'all')]),           # t3C3
    ('t3abb', [('v', 1, 'all'), ('v', 1, 'act'), ('v', 0, 'all'),
               ('o', 1, 'inact'), ('o', 1, 'inact'), ('o', 0, 'act')]),       # t3C4
    ('t3bbb', [('v', 1, 'all'), ('v', 1, 'all'), ('v', 1, 'act'),
               ('o', 1, 'all'), ('o', 1, 'all'), ('o', 1, 'act')]),           # t3D
]


def _range(act: ActiveSpace, space, spin, part, oa, ob, va, vb):
    if space == 'o':
        n = ob if spin else oa
        s = act.occ_start()
        return {'all': range(0, n), 'act': range(s, n), 'inact': range(0, s)}[part]
    n = vb if spin else va
    e = min(act.vir_end(spin), n)
    return {'all': range(0, n), 'act': range(0, e), 'inact': range(e, n)}[part]


def packed_index(act: ActiveSpace, oa, ob, va, vb):
    """[(block name, index arrays into the full C-order block)] in t_vec order."""
    out = []
    for name, slots in PACKED:
        ranges = [np.array(_range(act, *sl, oa, ob, va, vb), dtype=np.int64) for sl in slots]
        # Fortran order: slot 0 fastest -> build the grid with slot 5 outermost
        grids = np.meshgrid(*ranges[::-1], indexing='ij')   # grids[k] is slot 5-k
        # full-block axis a = 5 - slot s  ->  grids index k = 5 - s = a
        out.append((name, tuple(g.ravel() for g in grids)))
    return out


def packed_t3_size(act: ActiveSpace, oa, ob, va, vb):
    return sum(ix[0].size for _, ix in packed_index(act, oa, ob, va, vb))


def tvec(t, act: ActiveSpace, oa, ob, va, vb):
    head = [t[n].ravel() for n in S.TVEC_ORDER[:5]]
    body = [t[name][ix] for name, ix in packed_index(act, oa, ob, va, vb)]
    return np.concatenate(head + body)


def _images(name):
    """(axis permutation, parity) of every reordering inside the block's
    antisymmetric groups (same-spin occupied / virtual index runs)."""
    rank, nb = S.BLOCKS[name]
    na = rank - nb
    groups = [g for g in (list(range(0, na)), list(range(na, rank)),
                          list(range(rank, rank + na)), list(range(rank + na, 2 * rank)))
              if len(g) > 1]
    out = []
    for combo in itertools.product(*[list(itertools.permutations(g)) for g in groups]):
        perm = list(range(2 * rank))
        sign = 1
        for g, img in zip(groups, combo):
            for src, dst in zip(g, img):
                perm[src] = dst
            sign *= S.bf._parity([g.index(x) for x in img])
        out.append((perm, sign))
    return out


def set_tvec(t, vec, act: ActiveSpace, oa, ob, va, vb):
    """Unpack a t_vec: every stored T3 value goes to its tuple and, with the
    permutation sign, to all antisymmetric images (ccq's kernels read a P
    tuple through whichever packed block holds it)."""
    off = 0
    for n in S.TVEC_ORDER[:5]:
        sz = t[n].size
        t[n][...] = vec[off:off + sz].reshape(t[n].shape)
        off += sz
    for n in ('t3aaa', 't3aab', 't3abb', 't3bbb'):
        t[n][...] = 0.0
    for name, ix in packed_index(act, oa, ob, va, vb):
        sz = ix[0].size
        x = vec[off:off + sz]
        for perm, sign in _images(name):
            t[name][tuple(ix[perm[a]] for a in range(6))] = sign * x
        off += sz


def jacobi_iter(t, opts: S.Options, residual, fock_a, fock_b, Vaa, Vab, Vbb, act: ActiveSpace,
                log=None) -> S.Result:
    """src/solve_cc.f90:175-269 with update_clusters_t3_opt and the packed DIIS vector."""
    oa, ob = act.occ_a - act.froz, act.occ_b - act.froz
    va, vb = act.orbs - act.occ_a, act.orbs - act.occ_b
    D = S.block_denominators(t, fock_a, fock_b, oa, ob, opts.shift)
    n = opts.diis_space
    ring = [None] * (n + 1)
    conv = [1.0, 1.0, 1.0]
    en_cor = 0.0
    hist = []
    converged = failed = False
    it = 0
    for it in range(1, opts.max_iter + 1):
        sweep(t, residual, D, act)
        e_new = S.energy(fock_a, fock_b, Vaa, Vab, Vbb, oa, ob, t)
        de = en_cor - e_new
        conv = [conv[1], conv[2], de]
        en_cor = e_new
        slot = it % (n + 1)
        vec = tvec(t, act, oa, ob, va, vb)
        ring[slot - 1 if slot else n] = vec.copy()
        prev_slot = (it - 1) % (n + 1)
        prev = np.zeros_like(vec) if it == 1 else ring[prev_slot - 1 if prev_slot else n]
        res = math.sqrt(math.sqrt(float((prev - vec) @ (prev - vec))))
        if it % (n + 1) == 0:
            if log:
                log('      DIIS cycle')
            set_tvec(t, S.diis_extrapolate(ring, n), act, oa, ob, va, vb)
        hist.append((it, e_new, de, res))
        if log:
            log(S.fmt_iteration(it, e_new, de, res, 0.0))
        if all(abs(c) < opts.tol for c in conv):
            converged = True
            break
        if math.isnan(e_new):
            break
        if it == opts.max_iter:
            failed = True
    return S.Result(en_cor=en_cor, iterations=it, converged=converged, failed=failed,
                    history=hist, t=t)
