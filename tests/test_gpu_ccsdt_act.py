"""GPU parity for active-space CCSDt (ccq calc_type=CCSDt): HIP engine vs the
oracle of oracle/active.py (brute-force determinant residuals, ccq's
update_clusters_t3_opt order and masked T3 groups, packed ten-block t_vec).

Bar: one sweep to 1e-12 on ccq's packed t_vec; full solve energies per
iteration within 1e-9 and packed amplitudes within 1e-8; with every orbital
active the H4 CCSDT known answer (tests/test_small.py:44-47) to 1e-10.
"""

import json
import os
import sys

import numpy as np
import pytest

from oracle import active as A
from oracle import integrals as I
from oracle import solver as S
from ccq_amd import engine as E
from ccq_amd import solve as SV
from ccq_amd import synth
import gpu_helpers as G

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "h4_sto3g")
T3 = ('t3aaa', 't3aab', 't3abb', 't3bbb')


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _system(n, oa, ob, seed=11):
    s = synth.make_system(n, oa, 20, seed)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    return s, fa, fb, vaa, vab, vbb


def _engine(n, oa, ob, ao, au, ints, shift=0.0, diis=5):
    eng = E.Engine(0, oa, ob, n, "CCSDt", shift=shift, diis_space=diis, act_occ=ao, act_unocc=au)
    eng.set_ints(*ints)
    return eng


def _random_p(n, oa, ob, act, seed):
    rng = np.random.default_rng(seed)
    t = S.zero_amplitudes('CCSDT', oa, ob, n - oa, n - ob)
    for k in t:
        t[k] = G.antisym_random(rng, k, oa, ob, n - oa, n - ob, 0.05)
        if k in T3:
            t[k][~A.p_mask(k, t[k].shape, act)] = 0.0
    return t


@pytest.mark.parametrize("n,oa,ob,ao,au", [(6, 3, 3, 1, 1), (7, 3, 2, 1, 1), (7, 3, 2, 2, 2)])
def test_ccsdt_act_sweep_matches_oracle(n, oa, ob, ao, au):
    _say(f"case ccsdt-act sweep n={n} oa={oa} ob={ob} act={ao},{au}")
    s, fa, fb, vaa, vab, vbb = _system(n, oa, ob)
    act = A.ActiveSpace(0, oa, ob, n, ao, au)
    va, vb = n - oa, n - ob
    t = _random_p(n, oa, ob, act, 3)
    eng = _engine(n, oa, ob, ao, au, (fa, fb, vaa, vab, vbb), shift=0.03)
    vec = A.tvec(t, act, oa, ob, va, vb)
    assert eng.t_size == vec.size
    eng.set_t(vec)
    # packing round trip through the device (the random input is antisymmetric
    # to roundoff only, so images may differ from their stored value by ~1e-18)
    np.testing.assert_allclose(eng.get_t(), vec, atol=1e-15, rtol=0)
    eng.update()
    got = eng.get_t()
    prov = S.BruteForceResiduals(s.h, s.eri, n, oa, ob, 3)
    D = S.block_denominators(t, fa, fb, oa, ob, 0.03)
    A.sweep(t, prov, D, act)
    want = A.tvec(t, act, oa, ob, va, vb)
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=0)
    eng.close()


def test_ccsdt_act_full_solve_matches_oracle_trajectory():
    _say("case ccsdt-act full solve vs oracle")
    n, oa, ob, ao, au = 7, 3, 2, 1, 1
    s, fa, fb, vaa, vab, vbb = _system(n, oa, ob)
    act = A.ActiveSpace(0, oa, ob, n, ao, au)
    t = S.zero_amplitudes('CCSDT', oa, ob, n - oa, n - ob)
    opts = S.Options(calc_type='CCSDT', tol=1e-10, max_iter=60, diis_space=5)
    ref = A.jacobi_iter(t, opts, S.BruteForceResiduals(s.h, s.eri, n, oa, ob, 3),
                        fa, fb, vaa, vab, vbb, act)
    eng = _engine(n, oa, ob, ao, au, (fa, fb, vaa, vab, vbb))
    res = SV.jacobi_iter(eng, tol=1e-10, max_iter=60, diis_space=5, out=None)
    assert res.converged and ref.converged
    assert res.iterations == ref.iterations
    for a, b in zip(res.history, ref.history):
        assert abs(a[1] - b[1]) < 1e-9
        assert abs(a[3] - b[3]) < 1e-6
    want = A.tvec(t, act, oa, ob, n - oa, n - ob)
    assert np.abs(eng.get_t() - want).max() < 1e-8
    eng.close()


def test_h4_ccsdt_act_all_active_is_ccsdt_kat():
    """Every orbital active: CCSDt is CCSDT (ccq's H4 CCSDT total -1.978696408825)."""
    _say("case h4 ccsdt-act all active")
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    e1 = I.load_e1int(os.path.join(GOLD, "onebody.inp"), 4)
    e2, _ = I.load_e2int(os.path.join(GOLD, "twobody.inp"), 4)
    fa, fb = I.gen_fock(e1, e2, 2, 2)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    eng = _engine(4, 2, 2, 2, 2, (fa, fb, vaa, vab, vbb))
    res = SV.jacobi_iter(eng, tol=1e-11, max_iter=60, diis_space=5, out=None)
    assert res.converged
    assert abs(res.en_cor - (kat["CCSDT_total"] - kat["reference_energy"])) < 1e-10
    eng.close()


def test_ccsdt_act_requires_an_active_space():
    _say("case ccsdt-act config error")
    with pytest.raises(E.EngineError):
        E.Engine(0, 2, 2, 4, "CCSDt")
