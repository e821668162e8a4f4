"""GPU parity for CCSDT: HIP engine vs the brute-force determinant oracle,
and ccq's own H4 known answer (tests/test_small.py:44-47) reproduced on GPU.

Bar: one sweep bit-for-bit to 1e-12 against the oracle sweep in ccq's order;
converged correlation energy within 1e-10 of the published H4 CCSDT value.
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
from gpu_helpers import antisym_random

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "h4_sto3g")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,oa,ob,rhf", [(5, 2, 2, False), (5, 2, 2, True), (6, 3, 2, False),
                                          (6, 3, 3, False)])
def test_ccsdt_sweep_matches_bruteforce(n, oa, ob, rhf):
    _say(f"case ccsdt sweep n={n} oa={oa} ob={ob} rhf={rhf}")
    s = synth.make_system(n, oa, 20, 30 + n)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    va, vb = n - oa, n - ob
    rng = np.random.default_rng(n + oa)
    t = osolver.zero_amplitudes('CCSDT', oa, ob, va, vb)
    for k in t:
        t[k] = antisym_random(rng, k, oa, ob, va, vb, 0.05)
    if rhf:
        t['t2bb'] = t['t2aa'].copy()
        t['t1b'] = t['t1a'].copy()
        t['t3bbb'] = t['t3aaa'].copy()
        t['t3abb'] = osolver.tran3bto3c(t['t3aab'])
    eng = E.Engine(0, oa, ob, n, "CCSDT", rhf=rhf, shift=0.02)
    eng.set_ints(fa, fb, vaa, vab, vbb)
    eng.set_t(E.join_tvec(t, oa, ob, va, vb))
    eng.update()
    got = E.split_tvec(eng.get_t(), oa, ob, va, vb)
    prov = osolver.BruteForceResiduals(s.h, s.eri, n, oa, ob, 3)
    D = osolver.block_denominators(t, fa, fb, oa, ob, 0.02)
    osolver.sweep(t, osolver.sweep_order('CCSDT', rhf), prov, D)
    for k in t:
        np.testing.assert_allclose(got[k], t[k], atol=1e-12, rtol=0, err_msg=k)
    eng.close()


def test_h4_ccsdt_known_answer_on_gpu():
    """ccq's tests/h4_sto3g/CCSDT.inp (rhf, tol 1e-11): total -1.978696408825."""
    _say("case h4 ccsdt KAT")
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    e1 = I.load_e1int(os.path.join(GOLD, "onebody.inp"), 4)
    e2, enuc = I.load_e2int(os.path.join(GOLD, "twobody.inp"), 4)
    fa, fb = I.gen_fock(e1, e2, 2, 2)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    eng = E.Engine(0, 2, 2, 4, "CCSDT", rhf=True, diis_space=5)
    eng.set_ints(fa, fb, vaa, vab, vbb)
    res = S.jacobi_iter(eng, tol=1e-11, max_iter=60, diis_space=5, out=None)
    assert res.converged
    want = kat["CCSDT_total"] - kat["reference_energy"]
    assert abs(res.en_cor - want) < 1e-10
    eng.close()


def test_ccsdt_full_solve_matches_oracle_trajectory():
    """A 6-orbital open-shell CCSDT solve: every iteration's energy vs the oracle."""
    _say("case ccsdt full solve vs oracle")
    n, oa, ob = 6, 3, 2
    s = synth.make_system(n, oa, 20, 77)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    t = osolver.zero_amplitudes('CCSDT', oa, ob, n - oa, n - ob)
    opts = osolver.Options(calc_type='CCSDT', tol=1e-10, max_iter=60)
    ref = osolver.jacobi_iter(t, opts, osolver.BruteForceResiduals(s.h, s.eri, n, oa, ob, 3),
                              fa, fb, vaa, vab, vbb, oa, ob)
    eng = E.Engine(0, oa, ob, n, "CCSDT")
    eng.set_ints(fa, fb, vaa, vab, vbb)
    res = S.jacobi_iter(eng, tol=1e-10, max_iter=60, diis_space=5, out=None)
    assert res.converged and ref.converged
    assert res.iterations == ref.iterations
    for a, b in zip(res.history, ref.history):
        assert abs(a[1] - b[1]) < 1e-9
    got = E.split_tvec(eng.get_t(), oa, ob, n - oa, n - ob)
    for k in t:
        assert np.abs(got[k] - t[k]).max() < 1e-8, k
    eng.close()
