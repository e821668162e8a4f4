"""Mixed precision: FP32-MFMA contractions (mode 1: on FP64 storage; mode 2:
on FP32 storage of every tensor of the sweep) with FP64 refinement
(ccq_hip_set_precision, solve.Mixed), against the plain FP64 engine, which is
itself pinned to the oracle by the other GPU tests.

Bar: one FP32 sweep stays within 1e-5 of the FP64 sweep's change of the
amplitudes (relative) and differs from it (the FP32 path ran); a mixed solve
ends on the FP64 fixed point: energy within 1e-9 Eh and amplitudes within
1e-8 of the plain FP64 solve; FP32 graph replays equal eager FP32 launches;
irrep-blocked FP32-storage sweeps keep symmetry-forbidden amplitudes zero.
"""

import sys

import numpy as np
import pytest

from oracle import solver as osolver
from ccq_amd import engine as E
from ccq_amd import solve as S
from gpu_helpers import Case, antisym_random

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _random_state(c, method, seed, scale=0.05, rhf=False):
    rng = np.random.default_rng(seed)
    t = osolver.zero_amplitudes(method, c.ca, c.cb, c.va, c.vb)
    lvl = {"CCSD": 2, "CCSDT": 3}[method]
    for k in t:
        if osolver.BLOCKS[k][0] <= lvl:
            t[k] = antisym_random(rng, k, c.ca, c.cb, c.va, c.vb, scale)
    if rhf:
        # closed-shell relations of the beta blocks
        t["t1b"] = t["t1a"].copy()
        t["t2bb"] = t["t2aa"].copy()
        if lvl == 3:
            t["t3bbb"] = t["t3aaa"].copy()
            t["t3abb"] = osolver.tran3bto3c(t["t3aab"])
    return E.join_tvec(t, c.ca, c.cb, c.va, c.vb)


def _engine(c, method, **kw):
    eng = E.Engine(0, c.oa, c.ob, c.n, method, **kw)
    eng.set_ints(c.fa, c.fb, c.vaa, c.vab, c.vbb)
    return eng


@pytest.mark.parametrize("method,n,o,ob,rhf", [("CCSD", 24, 5, 5, False), ("CCSD", 24, 5, 5, True),
                                               ("CCSD", 30, 6, 4, False), ("CCSDT", 12, 3, 3, False),
                                               ("CCSDT", 12, 3, 3, True)])
@pytest.mark.parametrize("mode", [1, 2])
def test_fp32_sweep_close_to_fp64(method, n, o, ob, rhf, mode):
    _say(f"case fp32 sweep {method} n={n} o={o} ob={ob} rhf={rhf} mode={mode}")
    c = Case(n, o, 2 * n, 7, occ_b=ob)
    vec = _random_state(c, method, 5, rhf=rhf)
    out = []
    for prec in (0, mode):
        eng = _engine(c, method, shift=0.02, rhf=rhf)
        eng.set_t(vec)
        eng.set_precision(prec)
        eng.update()
        out.append(eng.get_t())
        eng.close()
    step = np.abs(out[0] - vec).max()
    diff = np.abs(out[1] - out[0]).max()
    assert diff > 1e-13, "the FP32 contractions did not run"
    assert diff < 1e-5 * step, (diff, step)


@pytest.mark.parametrize("method,n,o", [("CCSD", 24, 5), ("CCSDT", 12, 3)])
@pytest.mark.parametrize("mode", [1, 2])
def test_mixed_solve_reaches_fp64_fixed_point(method, n, o, mode):
    _say(f"case mixed solve {method} n={n} o={o} mode={mode}")
    c = Case(n, o, 2 * n, 11)
    plain = _engine(c, method)
    r0 = S.jacobi_iter(plain, tol=1e-10, max_iter=80, diis_space=5, out=None)
    mixed = _engine(c, method)
    r1 = S.jacobi_iter(mixed, tol=1e-10, max_iter=80, diis_space=5, out=None,
                       mixed=S.Mixed(switch_de=1e-6, mode=mode))
    assert r0.converged and r1.converged
    n32 = sum(1 for h in r1.history if h[5])
    assert n32 >= 3, "no FP32 sweeps were run"
    assert not r1.history[-1][5]
    assert abs(r1.en_cor - r0.en_cor) < 1e-9
    assert np.abs(mixed.get_t() - plain.get_t()).max() < 1e-8
    plain.close()
    mixed.close()


def test_mixed_refine_every_schedule():
    _say("case mixed refine_every")
    c = Case(20, 4, 40, 13)
    eng = _engine(c, "CCSD")
    r = S.jacobi_iter(eng, tol=1e-10, max_iter=80, diis_space=5, out=None,
                      mixed=S.Mixed(switch_de=1e-7, refine_every=3))
    assert r.converged
    flags = [h[5] for h in r.history]
    # every third sweep before the switch is FP64
    first64 = flags.index(False)
    assert first64 == 2
    ref = S.jacobi_iter(_engine(c, "CCSD"), tol=1e-10, max_iter=80, diis_space=5, out=None)
    assert abs(r.en_cor - ref.en_cor) < 1e-9
    eng.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_fp32_graph_replays_match_eager(mode, monkeypatch):
    _say(f"case fp32 graph replay, mode {mode}")
    c = Case(16, 4, 32, 17, occ_b=3)
    vec = _random_state(c, "CCSD", 9)
    engines = []
    for graph in ("1", "0"):
        monkeypatch.setenv("CCQ_HIP_GRAPH", graph)
        eng = _engine(c, "CCSD", shift=0.03)
        eng.set_t(vec)
        eng.set_precision(mode)
        engines.append(eng)
    for rep in range(3):
        out = [(eng.update(), eng.energy(), eng.get_t())[1:] for eng in engines]
        assert out[0][0] == out[1][0]
        np.testing.assert_array_equal(out[0][1], out[1][1])
    # switching back to FP64 uses the FP64 graph again
    for eng in engines:
        eng.set_precision(0)
    out = [(eng.update(), eng.get_t())[1] for eng in engines]
    np.testing.assert_array_equal(out[0], out[1])
    for eng in engines:
        eng.close()


def test_precision_mode_is_checked():
    _say("case precision mode error")
    c = Case(8, 3, 16, 1)
    eng = _engine(c, "CCSD")
    with pytest.raises(E.EngineError, match="precision"):
        eng.set_precision(3)
    with pytest.raises(E.EngineError, match="precision"):
        eng.set_precision(-1)
    eng.close()


def test_fp32_storage_blocked_sweep_and_new_integrals():
    """FP32 storage with irrep blocks (CCSDT, D2h): forbidden amplitudes stay
    exactly zero and the sweep stays within 1e-5 of the FP64 step; integrals
    uploaded again after an FP32 sweep reach the float shadows of the views
    (the next FP32 sweep matches a fresh engine on the new integrals)."""
    from oracle import integrals as I
    from ccq_amd import synth
    _say("case fp32 storage, irrep blocks, new integrals")
    n, o = 12, 3
    s = synth.make_system(n, o, 20, 81, d2h_irreps=True)
    fa, fb = I.gen_fock(s.h, s.e2int(), o, o)
    ints = (fa, fb) + tuple(I.antisymmetrize(s.e2int(), 0))
    out = []
    for prec in (0, 2):
        eng = E.Engine(0, o, o, n, "CCSDT", shift=0.02)
        eng.set_irreps(np.asarray(s.irreps) + 1)
        eng.set_ints(*ints)
        eng.set_precision(prec)
        for _ in range(3):
            eng.update()
        out.append(eng.get_t())
        if prec == 2:
            last = eng
        else:
            eng.close()
    step = np.abs(out[0]).max()
    diff = np.abs(out[1] - out[0]).max()
    assert 1e-13 < diff < 1e-5 * step, (diff, step)
    assert np.all(out[1][out[0] == 0.0] == 0.0)
    # new integrals (scaled two-electron part) on the engine that ran FP32
    ints2 = (fa, fb) + tuple(0.9 * v for v in ints[2:])
    start = out[1]
    last.set_ints(*ints2)
    last.set_t(start)
    last.update()
    got = last.get_t()
    last.close()
    fresh = E.Engine(0, o, o, n, "CCSDT", shift=0.02)
    fresh.set_irreps(np.asarray(s.irreps) + 1)
    fresh.set_ints(*ints2)
    fresh.set_precision(2)
    fresh.set_t(start)
    fresh.update()
    np.testing.assert_array_equal(got, fresh.get_t())
    fresh.close()


@pytest.mark.parametrize("method,rhf,kw", [("CCSDTQ", True, {}), ("CCSDTQ", False, {}),
                                           ("CCSDt", False, dict(act_occ=1, act_unocc=3))])
def test_fp32_storage_sweeps_from_zero(method, rhf, kw):
    """Three FP32-storage sweeps from T = 0 (closed-shell permuted copies of
    T3 and T4 into the amplitudes; active-space masked updates) stay within
    1e-5 of the FP64 sweeps (relative to the largest amplitude) and differ
    from them."""
    n, o = 9, 3
    _say(f"case fp32 storage sweeps {method} rhf={rhf} {kw}")
    c = Case(n, o, 2 * n, 23)
    out = []
    for prec in (0, 2):
        eng = _engine(c, method, shift=0.02, rhf=rhf, **kw)
        eng.set_precision(prec)
        for _ in range(3):
            eng.update()
        t = eng.get_t()
        if method == "CCSDTQ":
            t = np.concatenate([t, eng.get_t4()])
        out.append(t)
        eng.close()
    scale = np.abs(out[0]).max()
    diff = np.abs(out[1] - out[0]).max()
    assert 1e-13 < diff < 1e-5 * scale, (diff, scale)


def test_fp32_storage_copy_of_amplitudes_tracks_writes():
    """The float copy of the amplitudes in FP32-storage sweeps is refreshed
    only when something may have written them: a sweep after set_t sees the
    new amplitudes, and sweeps that follow sweeps (energy, DIIS record and
    residuum in between) equal sweeps after an explicit re-upload of the
    same amplitudes, bit for bit."""
    _say("case fp32 storage amplitude copy")
    c = Case(12, 3, 24, 19, occ_b=2)
    v1 = _random_state(c, "CCSDT", 41)
    v2 = _random_state(c, "CCSDT", 43)
    a = _engine(c, "CCSDT", shift=0.02)
    b = _engine(c, "CCSDT", shift=0.02)
    for eng in (a, b):
        eng.set_precision(2)
    a.set_t(v1)
    a.update()
    a.set_t(v2)
    a.update()
    b.set_t(v2)
    b.update()
    np.testing.assert_array_equal(a.get_t(), b.get_t())
    for it in range(1, 4):
        for eng, reload in ((a, False), (b, True)):
            if reload:
                eng.set_t(eng.get_t())
            eng.update()
            eng.energy()
            eng.diis_record(it)
            eng.residuum(it)
        np.testing.assert_array_equal(a.get_t(), b.get_t())
    a.close()
    b.close()
