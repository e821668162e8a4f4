"""Numpy evaluator of compiled block programs (CPU debugging of the IR).

Runs the same BlockOps the HIP executor runs, with ``numpy.einsum`` in place
of the GEMM kernels, so a compiled program can be checked on the CPU before
it is ever sent to a GPU.  Amplitude, integral and Fock arrays use the
layouts of the C-ABI (C order; amplitudes occupied-first like ccq's t_vec).
"""

from __future__ import annotations

import numpy as np

from .spinint import BlockKey, group_images

AMP_BLOCKS = {
    ('T1', (0, 0)): 't1a', ('T1', (1, 1)): 't1b',
    ('T2', (0, 0, 0, 0)): 't2aa', ('T2', (0, 1, 0, 1)): 't2ab', ('T2', (1, 1, 1, 1)): 't2bb',
    ('T3', (0, 0, 0, 0, 0, 0)): 't3aaa', ('T3', (0, 0, 1, 0, 0, 1)): 't3aab',
    ('T3', (0, 1, 1, 0, 1, 1)): 't3abb', ('T3', (1, 1, 1, 1, 1, 1)): 't3bbb',
    ('T4', (0, 0, 0, 0, 0, 0, 0, 0)): 't4aaaa', ('T4', (0, 0, 0, 1, 0, 0, 0, 1)): 't4aaab',
    ('T4', (0, 0, 1, 1, 0, 0, 1, 1)): 't4aabb', ('T4', (0, 1, 1, 1, 0, 1, 1, 1)): 't4abbb',
    ('T4', (1, 1, 1, 1, 1, 1, 1, 1)): 't4bbbb',
}
RES_OF = {v: ('R' + k[0][1:], k[1]) for k, v in AMP_BLOCKS.items()}
# left-CC amplitudes and residuals, CC(t;3) L3 / M3 accumulators
AMP_BLOCKS.update({
    ('L1', (0, 0)): 'l1a', ('L1', (1, 1)): 'l1b',
    ('L2', (0, 0, 0, 0)): 'l2aa', ('L2', (0, 1, 0, 1)): 'l2ab', ('L2', (1, 1, 1, 1)): 'l2bb',
})
for _k, _v in list(AMP_BLOCKS.items()):
    if _k[0] in ('L1', 'L2'):
        RES_OF[_v] = ('LR' + _k[0][1:], _k[1])
for _sp, _nm in (((0, 0, 0, 0, 0, 0), 'aaa'), ((0, 0, 1, 0, 0, 1), 'aab'),
                 ((0, 1, 1, 0, 1, 1), 'abb'), ((1, 1, 1, 1, 1, 1), 'bbb')):
    RES_OF['l3' + _nm] = ('L3', _sp)
    RES_OF['m3' + _nm] = ('M3', _sp)


class Dims:
    def __init__(self, oa, ob, va, vb):
        self.oa, self.ob, self.va, self.vb = oa, ob, va, vb
        assert oa + va == ob + vb
        self.n = oa + va

    def occ(self, spin):
        return self.oa if spin == 0 else self.ob

    def extent(self, space, spin):
        if space == 'o':
            return self.occ(spin)
        return self.n - self.occ(spin)


def base_view(key: BlockKey, dims: Dims, base: dict):
    """Slice of a full-range base array for a (spaces, spins) block."""
    if key.name == 'F':
        arr = base['Fa'] if key.spins[0] == 0 else base['Fb']
    elif key.name == 'V':
        sp = key.spins
        if sp == (0, 0, 0, 0):
            arr = base['Vaa']
        elif sp == (1, 1, 1, 1):
            arr = base['Vbb']
        elif sp == (0, 1, 0, 1):
            arr = base['Vab']
        else:
            raise KeyError(f"non-canonical V block {key}")
    else:
        raise KeyError(key)
    idx = tuple(slice(0, dims.occ(s)) if c == 'o' else slice(dims.occ(s), None)
                for c, s in zip(key.spaces, key.spins))
    return arr[idx]


def block_shape(key: BlockKey, dims: Dims):
    return tuple(dims.extent(c, s) for c, s in zip(key.spaces, key.spins))


class Interpreter:
    def __init__(self, prog, sections, dims: Dims, base: dict, amps: dict,
                 shift: float = 0.0, active=None):
        self.prog, self.sections, self.dims = prog, sections, dims
        self.base, self.amps = base, amps
        self.shift = shift
        # active space of masked updates: (first active occupied index,
        # active alpha virtuals, active beta virtuals)
        self.active = active
        self.store: dict[BlockKey, np.ndarray] = {}

    def get(self, key: BlockKey):
        kind = self.prog.decls[key.name].kind
        if kind == 'base':
            return base_view(key, self.dims, self.base)
        if kind == 'amp':
            return self.amps[AMP_BLOCKS[(key.name, key.spins)]]
        return self.store[key]

    def run_section(self, sec):
        if sec.kind == 'compute':
            for op in sec.ops:
                out_shape = block_shape(op.out, self.dims)
                terms = []
                spec = ",".join(l for _, l in op.ins) + "->" + op.out_labels
                val = np.einsum(spec, *[self.get(k) for k, _ in op.ins], optimize=True)
                val = float(op.coef) * val
                if self.prog.decls[op.out.name].kind == 'amp':
                    blk = self.amps[AMP_BLOCKS[(op.out.name, op.out.spins)]]
                    if op.assign:
                        blk[...] = val
                    else:
                        blk += val
                    continue
                if op.assign or op.out not in self.store:
                    self.store[op.out] = np.array(val, copy=True).reshape(out_shape)
                else:
                    self.store[op.out] += val
                del terms
        elif sec.kind == 'update':
            self.update(sec.extra['block'], sec.target, sec.extra.get('mask'))
        elif sec.kind == 'copy':
            self.amps[sec.extra['dst']][...] = self.amps[sec.extra['src']]

    def residual(self, block: str, key: BlockKey):
        U = self.store[key]
        decl = self.prog.decls[key.name]
        imgs = group_images(decl, key)
        R = np.zeros_like(U)
        for perm, sign in imgs:
            R += sign * U.transpose(perm)
        return R / len(imgs)

    def update(self, block: str, key: BlockKey, mask=None):
        R = self.residual(block, key)
        D = self.denominators(key)
        if mask is None:
            self.amps[block] -= R / D
            return
        m = self.mask(key, *mask)
        self.amps[block][m] -= R[m] / D[m]

    def mask(self, key: BlockKey, vneed: int, vforbid: int):
        """Tuples with an active occupied index, an active virtual of a spin in
        vneed and no active virtual of a spin in vforbid."""
        occ_start, vend_a, vend_b = self.active
        shape = block_shape(key, self.dims)
        idx = np.indices(shape)
        occ = np.zeros(shape, bool)
        vir = [np.zeros(shape, bool), np.zeros(shape, bool)]
        for p, (c, s) in enumerate(zip(key.spaces, key.spins)):
            if c == 'o':
                occ |= idx[p] >= occ_start
            else:
                vir[s] = vir[s] | (idx[p] < (vend_b if s else vend_a))
        m = np.zeros(shape, bool)
        for s in (0, 1):
            if vneed >> s & 1:
                m |= vir[s]
        m &= occ
        for s in (0, 1):
            if vforbid >> s & 1:
                m &= ~vir[s]
        return m

    def denominators(self, key: BlockKey):
        fa = np.diag(self.base['Fa'])
        fb = np.diag(self.base['Fb'])
        D = np.zeros(block_shape(key, self.dims))
        rank = len(key.spaces)
        for p, (c, s) in enumerate(zip(key.spaces, key.spins)):
            f = fa if s == 0 else fb
            o = self.dims.occ(s)
            e = f[:o] if c == 'o' else f[o:]
            shp = [1] * rank
            shp[p] = len(e)
            D = D + (e if c == 'v' else -e).reshape(shp)
        return D + self.shift

    def sweep(self):
        for sec in self.sections:
            self.run_section(sec)
