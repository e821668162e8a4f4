"""GPU parity for CCSDTQ: HIP engine vs the brute-force determinant oracle
(one sweep, live and through a stored fixture) and ccq's own H4 CCSDTQ known
answer (tests/test_small.py:38-52, rhf=.false., tol 1e-11).

Bar: one sweep to 1e-12 against the oracle sweep in ccq's order
(src/update_cc.F90:12-323, T4A..T4E after T3); converged correlation energy
within 1e-10 of the published H4 value.
"""

import json
import os
import sys

import numpy as np
import pytest

from oracle import integrals as I
from oracle import solver as osolver
from ccq_amd import engine as E
from ccq_amd import solve as S
from ccq_amd import synth
import gpu_helpers as G

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine_sweep(n, oa, ob, t, fa, fb, vaa, vab, vbb, shift, rhf=False):
    va, vb = n - oa, n - ob
    eng = E.Engine(0, oa, ob, n, "CCSDTQ", rhf=rhf, shift=shift)
    eng.set_ints(fa, fb, vaa, vab, vbb)
    eng.set_t(E.join_tvec(t, oa, ob, va, vb))
    eng.set_t4(E.join_t4(t, oa, ob, va, vb))
    eng.update()
    got = E.split_tvec(eng.get_t(), oa, ob, va, vb)
    got.update(E.split_t4(eng.get_t4(), oa, ob, va, vb))
    eng.close()
    return got


def test_ccsdtq_sweep_matches_stored_oracle_sweep():
    """8 orbitals, 4 alpha / 3 beta: every T4 block non-trivial (fixture made by
    golden/make_ccsdtq_sweep.py with the brute-force oracle)."""
    _say("case ccsdtq sweep n=8 oa=4 ob=3 vs fixture")
    c = G.ccsdtq_sweep_case()
    want = np.load(os.path.join(GOLD, "ccsdtq_sweep_n8.npz"))
    got = _engine_sweep(c["n"], c["oa"], c["ob"], c["t"], c["fa"], c["fb"], c["vaa"], c["vab"],
                        c["vbb"], c["shift"])
    for k in osolver.BLOCKS:
        x = got[k]
        np.testing.assert_allclose(x[G.canonical_mask(k, x.shape)], want[k], atol=1e-12, rtol=0,
                                   err_msg=k)
        # every image of a canonical tuple carries the signed value
        np.testing.assert_allclose(G.antisymmetrize_block(x, k), x, atol=1e-14, rtol=0, err_msg=k)


@pytest.mark.parametrize("rhf", [False, True])
def test_ccsdtq_sweep_matches_bruteforce(rhf):
    _say(f"case ccsdtq sweep n=6 oa=ob=3 rhf={rhf}")
    n, oa, ob = 6, 3, 3
    s = synth.make_system(n, oa, 20, 9)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    va, vb = n - oa, n - ob
    rng = np.random.default_rng(2)
    t = osolver.zero_amplitudes('CCSDTQ', oa, ob, va, vb)
    for k in t:
        t[k] = G.antisym_random(rng, k, oa, ob, va, vb, 0.03)
    if rhf:
        t['t2bb'] = t['t2aa'].copy()
        t['t1b'] = t['t1a'].copy()
        t['t3bbb'] = t['t3aaa'].copy()
        t['t3abb'] = osolver.tran3bto3c(t['t3aab'])
        t['t4abbb'] = osolver.tran4bto4d(t['t4aaab'])
        t['t4bbbb'] = t['t4aaaa'].copy()
    got = _engine_sweep(n, oa, ob, t, fa, fb, vaa, vab, vbb, 0.05, rhf=rhf)
    prov = osolver.BruteForceResiduals(s.h, s.eri, n, oa, ob, 4)
    D = osolver.block_denominators(t, fa, fb, oa, ob, 0.05)
    osolver.sweep(t, osolver.sweep_order('CCSDTQ', rhf), prov, D)
    for k in t:
        np.testing.assert_allclose(got[k], t[k], atol=1e-12, rtol=0, err_msg=k)


def test_h4_ccsdtq_known_answer_on_gpu():
    """ccq's tests/h4_sto3g/CCSDTQ.inp (rhf=.false., tol 1e-11): total -1.975861644132
    (4 electrons: CCSDTQ is FCI)."""
    _say("case h4 ccsdtq KAT")
    kat = json.load(open(os.path.join(GOLD, "h4_sto3g", "kat.json")))
    e1 = I.load_e1int(os.path.join(GOLD, "h4_sto3g", "onebody.inp"), 4)
    e2, enuc = I.load_e2int(os.path.join(GOLD, "h4_sto3g", "twobody.inp"), 4)
    fa, fb = I.gen_fock(e1, e2, 2, 2)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    eng = E.Engine(0, 2, 2, 4, "CCSDTQ", rhf=False, diis_space=5)
    eng.set_ints(fa, fb, vaa, vab, vbb)
    res = S.jacobi_iter(eng, tol=1e-11, max_iter=60, diis_space=5, out=None)
    assert res.converged
    want = kat["CCSDTQ_total"] - kat["reference_energy"]
    assert abs(res.en_cor - want) < 1e-10
    eng.close()
