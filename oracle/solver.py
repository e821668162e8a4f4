"""Restatement of ccq's CC iteration loop (oracle).

Follows src/solve_cc.f90 (jacobi_iter :175-269, convergence :293-316),
src/update_cc.F90 (update_clusters_t4 :12-323: in-place block sweep and RHF
copies), the Jacobi update loops of src/cc/t{1,2,3,4}*_update.f90
(t1A:518-525, t2A:1885-1901, t2B:2914-2926, t3A/t3B scatter loops),
src/energy.f90 (calculate_unsorted_energy :138-270), src/diis.F90 (ring,
residuum, B matrix with -1 border and corner, extrapolation over the older
slots) and src/printing.F90 (iteration and summary formats).

Amplitude blocks use the C-order layout of ccq's t_vec memory
(occupied indices, then virtual indices, alpha before beta).
"""

from __future__ import annotations

import dataclasses
import itertools
import math

import numpy as np

from . import bruteforce as bf
from . import ccsd as ccsd_mod

# block name -> (rank, number of beta index pairs)
BLOCKS = {
    't1a': (1, 0), 't1b': (1, 1),
    't2aa': (2, 0), 't2ab': (2, 1), 't2bb': (2, 2),
    't3aaa': (3, 0), 't3aab': (3, 1), 't3abb': (3, 2), 't3bbb': (3, 3),
    't4aaaa': (4, 0), 't4aaab': (4, 1), 't4aabb': (4, 2), 't4abbb': (4, 3), 't4bbbb': (4, 4),
}
TVEC_ORDER = ['t1a', 't1b', 't2aa', 't2ab', 't2bb', 't3aaa', 't3aab', 't3abb', 't3bbb']
LEVEL = {'CCSD': 2, 'CCSDT': 3, 'CCSDTQ': 4}


@dataclasses.dataclass
class Options:
    calc_type: str = 'CCSD'
    tol: float = 1e-4          # src/parser.f90:779-844 defaults
    diis_space: int = 5
    max_iter: int = 60
    shift: float = 0.0
    rhf: bool = False


def block_shape(name, oa, ob, va, vb):
    rank, nb = BLOCKS[name]
    return bf.block_shape(rank, nb, oa, ob, va, vb)


def zero_amplitudes(calc_type, oa, ob, va, vb):
    lvl = LEVEL[calc_type]
    out = {}
    for name, (rank, nb) in BLOCKS.items():
        # ccq allocates T1..T3 for every method (src/utils/cc_utils.f90:546-556)
        if rank <= max(lvl, 3):
            out[name] = np.zeros(block_shape(name, oa, ob, va, vb))
    return out


def denominators(name, fa_diag_o, fa_diag_v, fb_diag_o, fb_diag_v, shift):
    rank, nb = BLOCKS[name]
    na = rank - nb
    eo = [fa_diag_o] * na + [fb_diag_o] * nb
    ev = [fa_diag_v] * na + [fb_diag_v] * nb
    D = np.zeros([len(x) for x in eo] + [len(x) for x in ev])
    for k, e in enumerate(ev):
        shp = [1] * (2 * rank)
        shp[rank + k] = len(e)
        D = D + e.reshape(shp)
    for k, e in enumerate(eo):
        shp = [1] * (2 * rank)
        shp[k] = len(e)
        D = D - e.reshape(shp)
    return D + shift


def energy(fa, fb, Vaa, Vab, Vbb, oa, ob, t) -> float:
    """src/energy.f90:138-270 with bare f and antisymmetrised integrals."""
    o_a, v_a, o_b, v_b = slice(0, oa), slice(oa, None), slice(0, ob), slice(ob, None)
    t1a, t1b, t2aa, t2ab, t2bb = t['t1a'], t['t1b'], t['t2aa'], t['t2ab'], t['t2bb']
    e2a = 0.25 * np.einsum('efmn,mnef->', Vaa[v_a, v_a, o_a, o_a], t2aa)
    e2c = 0.25 * np.einsum('efmn,mnef->', Vbb[v_b, v_b, o_b, o_b], t2bb)
    e2b = np.einsum('efmn,mnef->', Vab[v_a, v_b, o_a, o_b], t2ab)
    e1a = np.einsum('me,me->', fa[o_a, v_a], t1a)
    e1b = np.einsum('me,me->', fb[o_b, v_b], t1b)
    e1a1a = 0.5 * np.einsum('efmn,nf,me->', Vaa[v_a, v_a, o_a, o_a], t1a, t1a)
    e1b1b = 0.5 * np.einsum('efmn,nf,me->', Vbb[v_b, v_b, o_b, o_b], t1b, t1b)
    e1a1b = np.einsum('efmn,me,nf->', Vab[v_a, v_b, o_a, o_b], t1a, t1b)
    return float(e1a + e1b + e2a + e2b + e2c + e1a1a + e1a1b + e1b1b)


def tran3bto3c(t3aab):
    """src/cc/tran3Bto3C.f90:1-31 -- closed-shell T3C from T3B."""
    return np.ascontiguousarray(t3aab.transpose(2, 0, 1, 5, 3, 4))


def tran4bto4d(t4aaab):
    """src/cc/tran4Bto4D.f:1-79 -- closed-shell T4D (abbb) from T4B (aaab)."""
    return np.ascontiguousarray(t4aaab.transpose(3, 0, 1, 2, 7, 4, 5, 6))


def sweep_order(calc_type: str, rhf: bool):
    """Block update order of src/update_cc.F90:102-318.

    Items are block names to update, or ('copy', dst, fn, src) for the RHF
    shortcuts (T2C = T2A, T1B = T1A, T3D = T3A, tran3bto3c, tran4bto4d, T4E).
    """
    lvl = LEVEL[calc_type]
    same = lambda x: x
    seq = ['t2aa']
    seq.append(('copy', 't2bb', same, 't2aa') if rhf else 't2bb')
    seq.append('t2ab')
    seq.append('t1a')
    seq.append(('copy', 't1b', same, 't1a') if rhf else 't1b')
    if lvl >= 3:
        seq.append('t3aaa')
        seq.append(('copy', 't3bbb', same, 't3aaa') if rhf else 't3bbb')
        seq.append('t3aab')
        seq.append(('copy', 't3abb', tran3bto3c, 't3aab') if rhf else 't3abb')
    if lvl >= 4:
        seq += ['t4aaaa', 't4aaab', 't4aabb']
        seq.append(('copy', 't4abbb', tran4bto4d, 't4aaab') if rhf else 't4abbb')
        # ccq's RHF T4E copy reads the T4A file through a REAL*4 buffer
        # (src/update_cc.F90:304-308); T4E is never read afterwards under rhf,
        # so the exact closed-shell copy is used here.
        seq.append(('copy', 't4bbbb', same, 't4aaaa') if rhf else 't4bbbb')
    return seq


class CCSDResiduals:
    """Residual provider: spin-integrated dressed CCSD (oracle/ccsd.py)."""

    def __init__(self, h, e2int, froz, occ_a, occ_b):
        self.h, self.e2int = h, e2int
        self.froz, self.occ_a, self.occ_b = froz, occ_a, occ_b
        self._key = None
        self._d = None
        self.ndress = 0

    def dressed(self, t):
        key = (t['t1a'].tobytes(), t['t1b'].tobytes())
        if key != self._key:
            self._d = ccsd_mod.dress(self.h, self.e2int, t['t1a'], t['t1b'],
                                     self.froz, self.occ_a, self.occ_b)
            self._key = key
            self.ndress += 1
        return self._d

    def __call__(self, name, t):
        d = self.dressed(t)
        fn = {'t2aa': ccsd_mod.r2aa, 't2bb': ccsd_mod.r2bb, 't2ab': ccsd_mod.r2ab,
              't1a': ccsd_mod.r1a, 't1b': ccsd_mod.r1b}[name]
        return fn(d, t)


class BruteForceResiduals:
    """Residual provider: determinant-space <mu|exp(-T)H exp(T)|0> (tiny only)."""

    def __init__(self, h, eri, norb, oa, ob, max_rank):
        self.space = bf.DetSpace(norb, oa, ob)
        self.H = bf.hamiltonian(self.space, h, eri)
        self.max_rank = max_rank
        self._key = None
        self._R = None

    def __call__(self, name, t):
        key = tuple(t[k].tobytes() for k in sorted(t))
        if key != self._key:
            amps = {}
            for rank in range(1, self.max_rank + 1):
                blocks = {BLOCKS[k][1]: v for k, v in t.items() if BLOCKS[k][0] == rank}
                amps[rank] = bf.blocks_to_so(self.space, rank, blocks)
            R, _ = bf.residuals(self.space, self.H, amps, range(1, self.max_rank + 1))
            self._R = {r: bf.so_to_blocks(self.space, r, R[r]) for r in R}
            self._key = key
        rank, nb = BLOCKS[name]
        return self._R[rank][nb]


def tvec(t, calc_type) -> np.ndarray:
    """ccq's t_vec prefix used by DIIS (src/solve_cc.f90:68-77)."""
    names = TVEC_ORDER[:5] if calc_type == 'CCSD' else TVEC_ORDER
    return np.concatenate([t[n].ravel() for n in names])


def set_tvec(t, calc_type, vec):
    names = TVEC_ORDER[:5] if calc_type == 'CCSD' else TVEC_ORDER
    off = 0
    for n in names:
        sz = t[n].size
        t[n][...] = vec[off:off + sz].reshape(t[n].shape)
        off += sz


def diis_extrapolate(ring, n):
    """src/diis.F90:11-84,159-242 (non-HDF5): slots 1..n+1 hold the last n+1 vectors."""
    B = np.zeros((n + 1, n + 1))
    for x in range(n):
        v3 = ring[x + 1] - ring[x]
        for y in range(x, n):
            B[x, y] += v3 @ (ring[y + 1] - ring[y])
    for x in range(n):
        for y in range(x + 1, n):
            B[y, x] = B[x, y]
    B[:, n] = -1.0
    B[n, :] = -1.0
    C = np.zeros(n + 1)
    C[n] = -1.0
    try:
        C = np.linalg.solve(B, C)
    except np.linalg.LinAlgError as exc:
        raise RuntimeError('calc_diis: DIIS error in DGESV.') from exc
    out = np.zeros_like(ring[0])
    for k in range(n):
        out += C[k] * ring[k]
    return out


def fmt_iteration(it, e, de, res, secs):
    nmin = int(secs) // 60
    nsec = secs - nmin * 60.0
    return f"  {it:4d}{e:15.10f}{de:15.10f}{res:15.10f}{nmin:5d} min{nsec:5.1f} s"


@dataclasses.dataclass
class Result:
    en_cor: float
    iterations: int
    converged: bool
    failed: bool
    history: list
    t: dict


def block_denominators(t, fock_a, fock_b, oa, ob, shift):
    fa_o, fa_v = np.diag(fock_a)[:oa], np.diag(fock_a)[oa:]
    fb_o, fb_v = np.diag(fock_b)[:ob], np.diag(fock_b)[ob:]
    return {name: denominators(name, fa_o, fa_v, fb_o, fb_v, shift) for name in t}


def sweep(t, seq, residual, D):
    """One in-place update_clusters_t4 sweep (src/update_cc.F90:102-318)."""
    for item in seq:
        if isinstance(item, tuple):
            _, dst, fn, src = item
            t[dst][...] = fn(t[src])
        else:
            R = residual(item, t)
            t[item] -= R / D[item]


def jacobi_iter(t, opts: Options, residual, fock_a, fock_b, Vaa, Vab, Vbb,
                oa, ob, log=None) -> Result:
    """src/solve_cc.f90:175-269.  ``t`` is updated in place."""
    D = block_denominators(t, fock_a, fock_b, oa, ob, opts.shift)
    seq = sweep_order(opts.calc_type, opts.rhf)
    n = opts.diis_space
    ring = [None] * (n + 1)
    conv = [1.0, 1.0, 1.0]           # src/solver_types.f90:22
    en_cor = 0.0                     # uninitialised in ccq (src/solver_types.f90:19)
    hist = []
    converged = failed = False
    it = 0
    for it in range(1, opts.max_iter + 1):
        sweep(t, seq, residual, D)
        e_new = energy(fock_a, fock_b, Vaa, Vab, Vbb, oa, ob, t)
        de = en_cor - e_new
        conv = [conv[1], conv[2], de]
        en_cor = e_new
        slot = it % (n + 1)
        vec = tvec(t, opts.calc_type)
        ring[slot - 1 if slot else n] = vec.copy()
        prev_slot = (it - 1) % (n + 1)
        prev = np.zeros_like(vec) if it == 1 else ring[prev_slot - 1 if prev_slot else n]
        res = math.sqrt(math.sqrt(float((prev - vec) @ (prev - vec))))
        if it % (n + 1) == 0:
            if log:
                log('      DIIS cycle')
            set_tvec(t, opts.calc_type, diis_extrapolate(ring, n))
        hist.append((it, e_new, de, res))
        if log:
            log(fmt_iteration(it, e_new, de, res, 0.0))
        if all(abs(c) < opts.tol for c in conv):
            converged = True
            break
        if math.isnan(e_new):
            if log:
                log('FAILED TO CONVERGE')
            break
        if it == opts.max_iter:
            failed = True
    return Result(en_cor=en_cor, iterations=it, converged=converged, failed=failed,
                  history=hist, t=t)


def unique_tuples(n, k):
    return itertools.combinations(range(n), k)
