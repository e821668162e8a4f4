"""The generated CC(t;3) tail program (ccq_amd/gen/cct3_program.py), run by the
numpy IR interpreter, against the oracle (oracle/cct3.py): H-bar blocks, the
left-CCSD residuals, l3 and the moments M3 for every spin block, at random
T1, T2, L1, L2.  This is the program embedded as ir_cct3_uhf.inc."""

import numpy as np
import pytest

from oracle import bruteforce as bf
from oracle import cct3 as C
from oracle import integrals as I
from oracle import solver as S
from ccq_amd import synth
from ccq_amd.gen import cct3_program as CP
from ccq_amd.gen import interp
from ccq_amd.gen.spinint import BlockKey
import gpu_helpers as G

L_BLOCKS = ('t1a', 't1b', 't2aa', 't2ab', 't2bb')


@pytest.fixture(scope="module")
def program():
    return CP.build()


@pytest.mark.parametrize("n,oa,ob", [(6, 3, 2), (6, 3, 3)])
def test_tail_program_matches_oracle(program, n, oa, ob):
    prog, secs = program
    s = synth.make_system(n, oa, 20, 5)
    space = bf.DetSpace(n, oa, ob)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    va, vb = n - oa, n - ob
    rng = np.random.default_rng(1)
    t = S.zero_amplitudes('CCSDT', oa, ob, va, vb)
    L = {}
    for k in L_BLOCKS:
        t[k] = G.antisym_random(rng, k, oa, ob, va, vb, 0.1)
        L['l' + k[1:]] = G.antisym_random(rng, k, oa, ob, va, vb, 0.1)
    amps = {k: v.copy() for k, v in t.items()}
    amps.update({k: v.copy() for k, v in L.items()})
    run = secs['hbar'] + secs['left'][:5] + [x for x in secs['cct3'] if x.kind != 'cct3']
    it = interp.Interpreter(prog, run, interp.Dims(oa, ob, va, vb),
                            dict(Fa=fa, Fb=fb, Vaa=vaa, Vab=vab, Vbb=vbb), amps)
    for sec in it.sections:
        it.run_section(sec)
    no = space.nocc_so
    so = lambda rank, blocks: bf.blocks_to_so(space, rank, blocks)
    t1 = so(1, {0: t['t1a'], 1: t['t1b']})
    t2 = so(2, {0: t['t2aa'], 1: t['t2ab'], 2: t['t2bb']})
    l1 = so(1, {0: L['l1a'], 1: L['l1b']})
    l2 = so(2, {0: L['l2aa'], 1: L['l2ab'], 2: L['l2bb']})
    hso, vso = C.so_integrals(space, s.h, s.eri)
    ht, gt = C.dress(hso, vso, t1)
    H = C.Hbar(ht, gt, t2, no)
    # one-body H-bar (the left denominators)
    sm = C.SpinMaps(n, oa, ob)
    for kind, mat in (('oo', H.oo), ('vv', H.vv)):
        for sp in (0, 1):
            cnt = (oa, ob)[sp] if kind == 'oo' else (va, vb)[sp]
            idx = [(sm.occ if kind == 'oo' else sm.vir)(sp, p) for p in range(cnt)]
            got = it.store[BlockKey('HB1', kind, (sp, sp))]
            np.testing.assert_allclose(got, mat[np.ix_(idx, idx)], atol=1e-14, rtol=0)
    # two-body H-bar blocks (the oracle keeps ovvo / vvvo: <ia||jb> = -<ia||bj>)
    two = dict(oooo=H.oooo, vvvv=H.vvvv, oovv=H.oovv, ooov=H.ooov, ovoo=H.ovoo,
               ovov=-H.ovvo.transpose(0, 1, 3, 2), vvov=-H.vvvo.transpose(0, 1, 3, 2),
               ovvv=-H.vovv.transpose(1, 0, 2, 3))
    checked = 0
    for key, got in it.store.items():
        if key.name != 'HB' or key.spaces not in two:
            continue
        idx = [[(sm.occ if c == 'o' else sm.vir)(s_, p) for p in range(n_)]
               for c, s_, n_ in zip(key.spaces, key.spins, got.shape)]
        np.testing.assert_allclose(got, two[key.spaces][np.ix_(*idx)], atol=1e-14, rtol=0)
        checked += 1
    assert checked >= 25
    # left residuals
    r1, r2 = C.lambda_residual(H, t2, l1, l2)
    b1, b2 = bf.so_to_blocks(space, 1, r1), bf.so_to_blocks(space, 2, r2)
    want = {'l1a': b1[0], 'l1b': b1[1], 'l2aa': b2[0], 'l2ab': b2[1], 'l2bb': b2[2]}
    for blk, ref in want.items():
        rn, sp = interp.RES_OF[blk]
        key = BlockKey(rn, 'o' * (len(sp) // 2) + 'v' * (len(sp) // 2), sp)
        np.testing.assert_allclose(it.residual(blk, key), ref, atol=1e-14, rtol=0)
    # l3 and the CCSD moments
    b3 = bf.so_to_blocks(space, 3, C.l3(H, l1, l2))
    R, _ = bf.residuals(space, bf.hamiltonian(space, s.h, s.eri), {1: t1, 2: t2}, [3])
    m3 = bf.so_to_blocks(space, 3, R[3])
    for nb, sfx in enumerate(('aaa', 'aab', 'abb', 'bbb')):
        for pre, ref in (('l3', b3), ('m3', m3)):
            rn, sp = interp.RES_OF[pre + sfx]
            key = BlockKey(rn, 'ooovvv', sp)
            got = it.residual(pre + sfx, key)
            np.testing.assert_allclose(got, ref[nb], atol=1e-14, rtol=0)
