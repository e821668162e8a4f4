"""GPU parity: HIP CCSD engine vs the CPU oracle, through the C ABI.

Bar: amplitudes within 1e-8 max-abs and energies within 1e-9 Eh of the
oracle (FP64 plain path); per-call quantities (one sweep, one energy, one
DIIS step) agree to ~1e-12.
"""

import math
import os
import sys

import numpy as np
import pytest

from oracle import solver as osolver
from ccq_amd import engine as E
from ccq_amd import solve as S
from gpu_helpers import Case

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,oa,ob,rhf", [(7, 3, 3, False), (7, 3, 3, True), (8, 3, 2, False),
                                          (9, 2, 2, False)])
def test_one_sweep_matches_oracle(n, oa, ob, rhf):
    _say(f"case sweep n={n} oa={oa} ob={ob} rhf={rhf}")
    c = Case(n, oa, 20, 100 + n, occ_b=ob)
    t = c.random_t(7)
    if rhf:
        t['t2bb'] = t['t2aa'].copy()
        t['t1b'] = t['t1a'].copy()
    eng = c.engine(rhf=rhf, shift=0.05)
    eng.set_t(c.tvec(t))
    # energy of the uploaded state
    e_gpu = eng.energy()
    e_ref = c.oracle_energy(t)
    assert abs(e_gpu - e_ref) < 1e-12
    eng.update()
    got = c.blocks(eng.get_t())
    D = osolver.block_denominators(t, c.fa, c.fb, c.ca, c.cb, 0.05)
    osolver.sweep(t, osolver.sweep_order('CCSD', rhf), c.provider(), D)
    for k in ['t1a', 't1b', 't2aa', 't2ab', 't2bb']:
        np.testing.assert_allclose(got[k], t[k], atol=1e-12, rtol=0, err_msg=k)
    eng.close()


def test_full_ccsd_n24_parity():
    """24 orbitals, 5 occupied, tol 1e-9, diis 5: energies and T at convergence."""
    _say("case config0 full solve")
    c = Case(24, 5, 48, 0)
    opts = osolver.Options(calc_type='CCSD', tol=1e-9, max_iter=60, diis_space=5)
    t = c.zero_t()
    ref = osolver.jacobi_iter(t, opts, c.provider(), c.fa, c.fb, c.vaa, c.vab, c.vbb, 5, 5)
    eng = c.engine()
    res = S.jacobi_iter(eng, tol=1e-9, max_iter=60, diis_space=5, out=None)
    assert res.converged and ref.converged
    assert res.iterations == ref.iterations
    for (it, e, de, r, *_), (it2, e2, de2, r2) in zip(res.history, ref.history):
        assert abs(e - e2) < 1e-9, (it, e, e2)
        assert abs(r - r2) < 1e-6
    assert abs(res.en_cor - ref.en_cor) < 1e-9
    got = c.blocks(eng.get_t())
    for k in ['t1a', 't1b', 't2aa', 't2ab', 't2bb']:
        assert np.abs(got[k] - t[k]).max() < 1e-8, k
    eng.close()


def test_ccsd_rhf_equals_uhf_at_convergence():
    _say("case rhf vs uhf")
    c = Case(12, 4, 24, 3)
    r1 = S.jacobi_iter(c.engine(rhf=False), 1e-10, 80, 5, out=None)
    r2 = S.jacobi_iter(c.engine(rhf=True), 1e-10, 80, 5, out=None)
    assert r1.converged and r2.converged
    assert abs(r1.en_cor - r2.en_cor) < 1e-9


def test_diis_and_residuum_match_oracle():
    _say("case diis")
    c = Case(8, 3, 16, 5)
    eng = c.engine(diis=3)
    rng = np.random.default_rng(3)
    vecs = []
    for it in range(1, 5):
        t = c.random_t(50 + it)
        eng.set_t(c.tvec(t))
        eng.diis_record(it)
        if it % 2 == 0:
            # a call in between: the residuum is recomputed from the ring
            # instead of taken from the record's fused pass
            eng.energy()
        r = eng.residuum(it)
        vec = osolver.tvec(t, 'CCSD')
        prev = vecs[-1] if vecs else np.zeros_like(vec)
        assert abs(r - math.sqrt(math.sqrt(float((prev - vec) @ (prev - vec))))) < 1e-12
        vecs.append(vec)
    eng.diis_extrapolate()
    got = osolver.tvec(c.blocks(eng.get_t()), 'CCSD')
    want = osolver.diis_extrapolate(vecs, 3)
    np.testing.assert_allclose(got, want, atol=1e-12, rtol=1e-10)
    del rng


def test_engine_error_paths():
    _say("case errors")
    c = Case(6, 2, 8, 1)
    eng = c.engine()
    with pytest.raises(E.EngineError, match="t_size"):
        eng.set_t(np.zeros(3))
    # identical DIIS vectors make the bordered system singular: ccq stops
    t = c.random_t(1)
    eng.set_t(c.tvec(t))
    for it in range(1, 7):
        eng.diis_record(it)
    with pytest.raises(E.EngineError, match="DGESV"):
        eng.diis_extrapolate()


@pytest.mark.parametrize("method", ["CCSD", "CCSDT"])
def test_graph_replays_match_eager(method, monkeypatch):
    """The captured sweep graph, replayed several times with eager work and
    synchronisations in between, equals the eagerly launched sweep each time."""
    _say(f"case graph replay {method}")
    from oracle import solver as osolver
    from gpu_helpers import antisym_random
    c = Case(7, 3, 20, 61, occ_b=2)
    rng = np.random.default_rng(4)
    t = osolver.zero_amplitudes(method, c.ca, c.cb, c.va, c.vb)
    lvl = {"CCSD": 2, "CCSDT": 3}[method]
    for k in t:
        if osolver.BLOCKS[k][0] <= lvl:
            t[k] = antisym_random(rng, k, c.ca, c.cb, c.va, c.vb, 0.05)
    vec = E.join_tvec(t, c.ca, c.cb, c.va, c.vb)
    engines = []
    for graph in ("1", "0"):
        monkeypatch.setenv("CCQ_HIP_GRAPH", graph)
        eng = E.Engine(0, c.oa, c.ob, c.n, method, shift=0.03)
        eng.set_ints(c.fa, c.fb, c.vaa, c.vab, c.vbb)
        eng.set_t(vec)
        engines.append(eng)
    for rep in range(3):
        out = []
        for eng in engines:
            eng.update()
            e = eng.energy()
            eng.diis_record(rep + 1)
            out.append((e, eng.get_t()))
        assert abs(out[0][0] - out[1][0]) < 1e-13
        np.testing.assert_array_equal(out[0][1], out[1][1])
    for eng in engines:
        eng.close()
