"""Point-group blocking (ccq_hip_set_irreps) against the dense computation.

ccq reads the orbital irreps (src/symmetry.f90:84-114) and still contracts
dense blocks.  With the irreps given, the engine splits every product into
its irrep blocks (a totally symmetric A[m,k] vanishes unless
irrep(m) = irrep(k)), so symmetry-forbidden elements are never computed.
On systems with D2h irreps and forbidden integrals zeroed, sweeps, the
CCSDt P-space sweep and the CC(t;3) tail must give the dense results
(1e-11), forbidden amplitudes must stay exactly zero, and the blocked sweep
must do fewer flops.
"""

import sys

import numpy as np
import pytest

from oracle import active as A
from oracle import integrals as I
from oracle import solver as S
from ccq_amd import engine as E
from ccq_amd import solve as SV
from ccq_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _system(n, o, seed):
    s = synth.make_system(n, o, 20, seed, d2h_irreps=True)
    fa, fb = I.gen_fock(s.h, s.e2int(), o, o)
    vaa, vab, vbb = I.antisymmetrize(s.e2int(), 0)
    return s, (fa, fb, vaa, vab, vbb)


def _engine(s, ints, method, sym, monkeypatch, graph="1", **kw):
    monkeypatch.setenv("CCQ_HIP_GRAPH", graph)
    eng = E.Engine(0, s.nocc, s.nocc, s.norb, method, **kw)
    if sym:
        eng.set_irreps(np.asarray(s.irreps) + 1)
    eng.set_ints(*ints)
    return eng


def _forbidden(name, shape, irreps, o):
    rank = len(shape) // 2
    g = np.zeros(shape, dtype=np.int64)
    ix = np.indices(shape)
    for ax in range(2 * rank):
        orb = ix[ax] if ax < rank else o + ix[ax]
        g ^= irreps[orb]
    return g != 0


def _sweeps(eng, count):
    es = []
    for _ in range(count):
        eng.update()
        es.append(eng.energy())
    return np.array(es)


def _flops(eng):
    eng.profile(True)
    eng.update()
    fl = sum(v["flops"] for v in eng.profile_report().values())
    eng.profile(False)
    return fl


@pytest.mark.parametrize("method,n,o,rhf", [("CCSD", 14, 4, False), ("CCSD", 14, 4, True),
                                            ("CCSDT", 11, 3, False), ("CCSDT", 11, 3, True),
                                            ("CCSDTQ", 9, 3, False)])
def test_blocked_sweeps_equal_dense(method, n, o, rhf, monkeypatch):
    _say(f"case irrep blocks {method} n={n} o={o} rhf={rhf}")
    s, ints = _system(n, o, 70 + n)
    v = n - o
    out = []
    for sym in (False, True):
        eng = _engine(s, ints, method, sym, monkeypatch, rhf=rhf, shift=0.02)
        es = _sweeps(eng, 4)
        t = E.split_tvec(eng.get_t(), o, o, v, v)
        t4 = E.split_t4(eng.get_t4(), o, o, v, v) if method == "CCSDTQ" else None
        eng.set_t(np.zeros(eng.t_size))
        if eng.t4_size:
            eng.set_t4(np.zeros(eng.t4_size))
        fl = _flops(eng)
        eng.close()
        out.append((es, t, t4, fl))
    (e0, t0, q0, f0), (e1, t1, q1, f1) = out
    _say(f"  sweep flops dense {f0 / 1e6:.1f} M, blocked {f1 / 1e6:.1f} M")
    np.testing.assert_allclose(e1, e0, rtol=1e-11, atol=0)
    for k in t0:
        if np.abs(t0[k]).max() == 0:
            continue
        assert np.abs(t1[k] - t0[k]).max() <= 1e-11 * np.abs(t0[k]).max(), k
        forb = _forbidden(k, t1[k].shape, s.irreps, o)
        assert np.all(t1[k][forb] == 0.0), k
    if q0 is not None:
        for k in q0:
            assert np.abs(q1[k] - q0[k]).max() <= 1e-11 * max(np.abs(q0[k]).max(), 1e-12), k
            assert np.all(q1[k][_forbidden(k, q1[k].shape, s.irreps, o)] == 0.0), k
    assert f1 < 0.5 * f0


def test_blocked_ccsdt_act_and_cct3_tail_equal_dense(monkeypatch):
    """CCSDt (P-space products and irrep blocks together), then H-bar, the
    left-CCSD solve and crcc23's four energies, blocked against dense."""
    n, o, ao, au = 11, 3, 1, 3
    _say(f"case irrep blocks CC(t;3) n={n} o={o} act {ao}/{au}")
    s, ints = _system(n, o, 91)
    v = n - o
    act = A.ActiveSpace(0, o, o, n, ao, au)
    out = []
    for sym in (False, True):
        eng = _engine(s, ints, "CC(t;3)", sym, monkeypatch, act_occ=ao, act_unocc=au)
        res = SV.jacobi_iter(eng, tol=1e-10, max_iter=60, diis_space=5, out=None)
        assert res.converged
        t = S.zero_amplitudes('CCSDT', o, o, v, v)
        A.set_tvec(t, eng.get_t(), act, o, o, v, v)
        eng.hbar()
        left = SV.solve_lcc(eng, res.en_cor, 1e-10, 60, 5, out=None)
        e23 = np.array(eng.cct3())
        eng.close()
        out.append((res.en_cor, res.iterations, t, left.iterations, e23))
        _say(f"  sym={sym}: {res.iterations} CCSDt iterations, E = {res.en_cor:.12f}, "
             f"left {left.iterations}, CC(t;3) {e23}")
    (ea, ia, ta, la, da), (eb, ib, tb, lb, db) = out
    assert ia == ib and la == lb
    assert abs(ea - eb) < 1e-11
    np.testing.assert_allclose(db, da, rtol=1e-10, atol=1e-13)
    for k in ta:
        assert np.abs(tb[k] - ta[k]).max() <= 1e-10 * max(np.abs(ta[k]).max(), 1e-12), k
        assert np.all(tb[k][_forbidden(k, tb[k].shape, s.irreps, o)] == 0.0), k


def test_irreps_must_precede_integrals(monkeypatch):
    s, ints = _system(8, 2, 5)
    eng = _engine(s, ints, "CCSD", False, monkeypatch)
    with pytest.raises(E.EngineError, match="before ccq_hip_set_ints"):
        eng.set_irreps(np.asarray(s.irreps) + 1)
    eng.close()
    eng = E.Engine(0, 2, 2, 8, "CCSD")
    with pytest.raises(E.EngineError, match="1..8"):
        eng.set_irreps(np.zeros(8, dtype=np.int64))
    eng.close()
