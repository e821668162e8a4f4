"""GPU parity for the CC(t;3) tail (ccq calc_type=CCT3): HIP engine vs the
oracle of oracle/cct3.py on converged active-space CCSDt amplitudes stored in
tests/golden/cct3_n7_*.npz (made by tests/golden/make_cct3.py).

Bar: H-bar blocks to 1e-13; the left-CCSD iteration (update_l2 Jacobi with
H-bar denominators, DIIS on l_vec) with the oracle's pseudo-energies per
iteration within 1e-9, the same iteration count and l_vec within 1e-8;
crcc23's four energies within 1e-11 Eh; end to end (CCSDt from zero, H-bar,
left-CCSD, crcc23) CC(t;3) energies within 1e-9 Eh.
"""

import os
import sys

import numpy as np
import pytest

from oracle import active as A
from oracle import bruteforce as bf
from oracle import cct3 as C
from oracle import integrals as I
from oracle import solver as S
from ccq_amd import engine as E
from ccq_amd import solve as SV
from ccq_amd import synth

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
N, AO, AU, SEED = 7, 1, 1, 11
CASES = [(3, 3), (3, 2)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(oa, ob):
    g = np.load(os.path.join(GOLD, f"cct3_n7_{oa}{ob}.npz"))
    s = synth.make_system(N, oa, 20, SEED)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    eng = E.Engine(0, oa, ob, N, "CC(t;3)", diis_space=5, act_occ=AO, act_unocc=AU)
    eng.set_ints(fa, fb, vaa, vab, vbb)
    return g, s, eng


def _oracle_hbar(s, oa, ob, tvec):
    act = A.ActiveSpace(0, oa, ob, N, AO, AU)
    va, vb = N - oa, N - ob
    t = S.zero_amplitudes('CCSDT', oa, ob, va, vb)
    A.set_tvec(t, tvec, act, oa, ob, va, vb)
    space = bf.DetSpace(N, oa, ob)
    t1 = bf.blocks_to_so(space, 1, {0: t['t1a'], 1: t['t1b']})
    t2 = bf.blocks_to_so(space, 2, {0: t['t2aa'], 1: t['t2ab'], 2: t['t2bb']})
    hso, vso = C.so_integrals(space, s.h, s.eri)
    ht, gt = C.dress(hso, vso, t1)
    return C.Hbar(ht, gt, t2, space.nocc_so), C.SpinMaps(N, oa, ob)


def _block(H, sm, name, spaces, spins):
    """Spin block of an H-bar block (e.g. 'ovov', spins (1, 0, 1, 0))."""
    full = getattr(H, name)
    idx = []
    for c, s in zip(spaces, spins):
        cnt = (sm.oa, sm.ob)[s] if c == 'o' else (sm.va, sm.vb)[s]
        idx.append([(sm.occ if c == 'o' else sm.vir)(s, p) for p in range(cnt)])
    return full[np.ix_(*idx)]


SPIN = {'a': 0, 'b': 1}
HB_BLOCKS = ['HB1.ov.aa', 'HB1.ov.bb', 'HB1.oo.aa', 'HB1.oo.bb', 'HB1.vv.aa', 'HB1.vv.bb',
             'HB.oooo.aaaa', 'HB.oooo.abab', 'HB.oooo.bbbb', 'HB.vvvv.aaaa', 'HB.vvvv.abab',
             'HB.vvvv.bbbb', 'HB.ovov.aaaa', 'HB.ovov.baba', 'HB.ovov.abab', 'HB.ovov.bbbb',
             'HB.oovv.aaaa', 'HB.oovv.abab', 'HB.oovv.bbbb', 'HB.ooov.aaaa', 'HB.ooov.abab',
             'HB.ovoo.aaaa', 'HB.ovoo.abab', 'HB.vvov.aaaa', 'HB.vvov.abab']


@pytest.mark.parametrize("oa,ob", CASES)
def test_hbar_blocks_match_oracle(oa, ob):
    _say(f"case cct3 hbar oa={oa} ob={ob}")
    g, s, eng = _setup(oa, ob)
    eng.set_t(g['tvec'])
    eng.hbar()
    H, sm = _oracle_hbar(s, oa, ob, g['tvec'])
    # the oracle keeps ovvo; <ia||jb> = -<ia||bj>
    H.ovov = -H.ovvo.transpose(0, 1, 3, 2)
    H.vvov = -H.vvvo.transpose(0, 1, 3, 2)
    for tid in HB_BLOCKS:
        _, name, sp = tid.split('.')
        spins = tuple(SPIN[c] for c in sp)
        want = _block(H, sm, name, name, spins)
        got = eng.debug_read('cct3:' + tid, want.shape)
        assert np.abs(got - want).max() < 1e-13, tid
    # the left denominators: [occ | vir] diagonals of the one-body blocks
    for sp, c in ((0, 'a'), (1, 'b')):
        want = np.concatenate([np.diag(_block(H, sm, 'oo', 'oo', (sp, sp))),
                               np.diag(_block(H, sm, 'vv', 'vv', (sp, sp)))])
        got = eng.debug_read('cct3:HD.' + c, (N,))
        assert np.abs(got - want).max() < 1e-13
    eng.close()


@pytest.mark.parametrize("oa,ob", CASES)
def test_left_solve_and_correction_match_oracle(oa, ob):
    _say(f"case cct3 left+crcc23 oa={oa} ob={ob}")
    g, s, eng = _setup(oa, ob)
    eng.set_t(g['tvec'])
    assert abs(eng.energy() - float(g['ecor'])) < 1e-13
    eng.hbar()
    res = SV.solve_lcc(eng, float(g['ecor']), tol=1e-11, max_iter=60, diis_space=5, out=None)
    assert res.converged
    want = g['left_e']
    assert res.iterations == len(want)
    for (it, e, *_), w in zip(res.history, want):
        assert abs(e - w) < 1e-9, it
    assert np.abs(eng.get_l() - g['l_vec']).max() < 1e-8
    e23 = eng.cct3()
    assert np.abs(np.array(e23) - g['e23']).max() < 1e-11
    # the same correction from the oracle's L
    eng.set_l(g['l_vec'])
    e23 = eng.cct3()
    assert np.abs(np.array(e23) - g['e23']).max() < 1e-12
    # T is untouched by the tail (the stored vector is antisymmetric to roundoff,
    # so a pack / unpack round trip may move an image by ~1e-20)
    np.testing.assert_allclose(eng.get_t(), g['tvec'], atol=1e-15, rtol=0)
    eng.close()


@pytest.mark.parametrize("oa,ob", CASES)
def test_cct3_end_to_end(oa, ob):
    _say(f"case cct3 end to end oa={oa} ob={ob}")
    g, s, eng = _setup(oa, ob)
    r = SV.cct3(eng, tol=1e-11, max_iter=80, diis_space=5, out=None)
    assert r.ccsdt.converged and r.left.converged
    assert abs(r.ccsdt.en_cor - float(g['ecor'])) < 1e-9
    assert np.abs(np.array(r.e23) - g['e23']).max() < 1e-10
    assert abs(r.en_cor - (float(g['ecor']) + g['e23'][3])) < 1e-9
    text = SV.print_cct3(-1.0, r.ccsdt.en_cor, r.e23)
    assert "CC(t;3),A" in text and text.splitlines()[-1].split()[0] == "CC(t;3)"
    eng.close()


def test_tail_entry_points_need_level_6():
    _say("case cct3 level check")
    eng = E.Engine(0, 3, 3, N, "CCSDt", act_occ=1, act_unocc=1)
    with pytest.raises(E.EngineError):
        eng.hbar()
    with pytest.raises(E.EngineError):
        eng.left_mode(True)
    assert eng.l_size == 0
    eng.close()
