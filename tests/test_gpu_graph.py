"""The captured sweep graph gives bit-for-bit the amplitudes of eager
launches, replay after replay, with eager work and synchronisations in
between -- for CCSD, CCSDT (FP64, FP32 contractions, FP32 storage), CCSDTQ,
active-space CCSDt and the left-CC program of CC(t;3)."""

import sys

import numpy as np
import pytest

from oracle import solver as osolver
from ccq_amd import engine as E
from gpu_helpers import Case, antisym_random

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _state(c, method, seed, scale=0.05):
    rng = np.random.default_rng(seed)
    t = osolver.zero_amplitudes("CCSDT", c.ca, c.cb, c.va, c.vb)
    lvl = {"CCSD": 2, "CCSDT": 3, "CCSDTQ": 3, "CCSDt": 3, "CC(t;3)": 3}[method]
    for k in t:
        if osolver.BLOCKS[k][0] <= lvl:
            t[k] = antisym_random(rng, k, c.ca, c.cb, c.va, c.vb, scale)
    return t


def _engines(c, method, monkeypatch, **kw):
    out = []
    for graph in ("0", "1"):
        monkeypatch.setenv("CCQ_HIP_GRAPH", graph)
        eng = E.Engine(0, c.oa, c.ob, c.n, method, **kw)
        eng.set_ints(c.fa, c.fb, c.vaa, c.vab, c.vbb)
        out.append(eng)
    return out


@pytest.mark.parametrize("method,n,o,ob,prec", [("CCSD", 16, 4, 3, 0), ("CCSDT", 9, 3, 2, 0),
                                                ("CCSDT", 9, 3, 2, 1), ("CCSDT", 9, 3, 2, 2),
                                                ("CCSDTQ", 8, 3, 2, 0), ("CCSDTQ", 8, 3, 2, 2),
                                                ("CCSDt", 9, 3, 2, 0), ("CCSDt", 9, 3, 2, 2)])
def test_graph_equals_eager(method, n, o, ob, prec, monkeypatch):
    _say(f"case graph {method} n={n} prec={prec}")
    c = Case(n, o, 2 * n, 31, occ_b=ob)
    t = _state(c, method, 5)
    kw = dict(shift=0.03)
    if method == "CCSDt":
        kw.update(act_occ=1, act_unocc=2)
    engines = _engines(c, method, monkeypatch, **kw)
    if method == "CCSDt":
        from oracle import active as A
        act = A.ActiveSpace(0, c.oa, c.ob, c.n, 1, 2)
        for k in ("t3aaa", "t3aab", "t3abb", "t3bbb"):
            t[k][~A.p_mask(k, t[k].shape, act)] = 0.0
        vec = A.tvec(t, act, c.ca, c.cb, c.va, c.vb)
    else:
        if method == "CCSD":
            t = {k: v for k, v in t.items() if osolver.BLOCKS[k][0] <= 2}
            t.update({k: np.zeros(osolver.block_shape(k, c.ca, c.cb, c.va, c.vb))
                      for k in ("t3aaa", "t3aab", "t3abb", "t3bbb")})
        vec = E.join_tvec(t, c.ca, c.cb, c.va, c.vb)
    for eng in engines:
        eng.set_t(vec)
        eng.set_precision(prec)
    for rep in range(3):
        out = []
        for eng in engines:
            eng.update()
            e = eng.energy()
            eng.diis_record(rep + 1)
            out.append((e, eng.get_t(), eng.get_t4() if method == "CCSDTQ" else None))
        for e, tv, t4 in out[1:]:
            assert e == out[0][0]
            np.testing.assert_array_equal(tv, out[0][1])
            if t4 is not None:
                np.testing.assert_array_equal(t4, out[0][2])
    for eng in engines:
        eng.close()


def test_left_graph_equals_eager(monkeypatch):
    _say("case graph left-CC")
    c = Case(9, 3, 18, 41)
    t = _state(c, "CC(t;3)", 7)
    from oracle import active as A
    act = A.ActiveSpace(0, c.oa, c.ob, c.n, 1, 1)
    for k in ("t3aaa", "t3aab", "t3abb", "t3bbb"):
        t[k][~A.p_mask(k, t[k].shape, act)] = 0.0
    vec = A.tvec(t, act, c.ca, c.cb, c.va, c.vb)
    engines = _engines(c, "CC(t;3)", monkeypatch, act_occ=1, act_unocc=1)
    for eng in engines:
        eng.set_t(vec)
        eng.hbar()
        eng.left_init(-0.1)
    for rep in range(3):
        out = []
        for eng in engines:
            eng.left_update()
            out.append((eng.left_energy(), eng.get_l()))
        for e, lv in out[1:]:
            assert e == out[0][0]
            np.testing.assert_array_equal(lv, out[0][1])
    for eng in engines:
        eng.close()
