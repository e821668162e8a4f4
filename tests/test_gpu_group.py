"""Grouped GEMM launches (independent tile GEMMs of one dependency level run
as one launch at full K) against one launch per GEMM in program order
(CCQ_HIP_GROUP=0), on systems large enough that groups form.

Bar: the profile shows grouped classes; energies after three sweeps agree to
1e-11 relative and amplitudes to 1e-11 of their largest entry (the two paths
differ only by the split-K summation order of the ungrouped launches); the
captured graph of the grouped sweep replays bit for bit as eager launches.
"""

import re
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


def _state(c, method, seed, scale=0.02):
    rng = np.random.default_rng(seed)
    t = osolver.zero_amplitudes(method, c.ca, c.cb, c.va, c.vb)
    lvl = {"CCSD": 2, "CCSDT": 3}[method]
    for k in t:
        if osolver.BLOCKS[k][0] <= lvl:
            t[k] = antisym_random(rng, k, c.ca, c.cb, c.va, c.vb, scale)
    return E.join_tvec(t, c.ca, c.cb, c.va, c.vb)


def _engine(c, method, monkeypatch, group, graph="0"):
    monkeypatch.setenv("CCQ_HIP_GROUP", group)
    monkeypatch.setenv("CCQ_HIP_GRAPH", graph)
    eng = E.Engine(0, c.oa, c.ob, c.n, method, shift=0.02)
    eng.set_ints(c.fa, c.fb, c.vaa, c.vab, c.vbb)
    return eng


GROUPED = re.compile(r"^gemm:.*\*.*g$")
# one launch of several flat block copies of a level (axpy_group)
ADD_GROUP = re.compile(r"^add:flat\*(\d+)$")


@pytest.mark.parametrize("method,n,o,ob", [("CCSD", 64, 8, 8), ("CCSD", 72, 9, 8), ("CCSDT", 30, 6, 6)])
def test_grouped_launches_match_single(method, n, o, ob, monkeypatch):
    _say(f"case grouped launches {method} n={n} o={o} ob={ob}")
    c = Case(n, o, 2 * n, 23, occ_b=ob)
    vec = _state(c, method, 3)
    out = []
    for group in ("0", "1"):
        eng = _engine(c, method, monkeypatch, group)
        eng.set_t(vec)
        eng.profile(True)
        eng.update()
        rep = eng.profile_report()
        eng.profile(False)
        grouped = [k for k in rep if GROUPED.match(k)]
        add_groups = [k for k in rep if ADD_GROUP.match(k)]
        if group == "1":
            _say("  grouped classes: " + ", ".join(grouped[:6] + add_groups[:4]))
            assert grouped, sorted(rep)[:20]
            assert add_groups, sorted(rep)[:40]
            assert all(2 <= int(ADD_GROUP.match(k).group(1)) <= 16 for k in add_groups)
        else:
            assert not grouped and not add_groups
        es = [eng.energy()]
        for _ in range(2):
            eng.update()
            es.append(eng.energy())
        out.append((np.array(es), eng.get_t()))
        eng.close()
    (e0, t0), (e1, t1) = out
    np.testing.assert_allclose(e1, e0, rtol=1e-11, atol=0)
    assert np.abs(t1 - t0).max() <= 1e-11 * np.abs(t0).max()


def test_grouped_graph_replays_equal_eager(monkeypatch):
    _say("case grouped graph replay")
    c = Case(64, 8, 128, 29)
    vec = _state(c, "CCSD", 5)
    engines = [_engine(c, "CCSD", monkeypatch, "1", graph) for graph in ("1", "0")]
    for eng in engines:
        eng.set_t(vec)
    for rep in range(3):
        res = []
        for eng in engines:
            eng.update()
            e = eng.energy()
            eng.diis_record(rep + 1)
            res.append((e, eng.get_t()))
        assert res[0][0] == res[1][0]
        np.testing.assert_array_equal(res[0][1], res[1][1])
    for eng in engines:
        eng.close()


def _engine_env(c, method, monkeypatch, env, **kw):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    eng = E.Engine(0, c.oa, c.ob, c.n, method, **kw)
    eng.set_ints(c.fa, c.fb, c.vaa, c.vab, c.vbb)
    return eng


@pytest.mark.parametrize("method,n,o,ob", [("CCSD", 16, 4, 3), ("CCSD", 24, 5, 5), ("CCSDT", 10, 3, 2),
                                           ("CCSDTQ", 8, 3, 2), ("CCSDt", 10, 3, 2)])
def test_packed_pairs_match_full(method, n, o, ob, monkeypatch):
    """Antisymmetric pairs packed to their canonical half (summed pairs
    doubled, output pairs mirrored with the opposite sign) against every
    element computed (CCQ_HIP_PACK=0): three sweeps agree to 1e-11."""
    _say(f"case packed pairs {method} n={n} o={o} ob={ob}")
    c = Case(n, o, 2 * n, 37, occ_b=ob)
    rng = np.random.default_rng(11)
    t = osolver.zero_amplitudes("CCSDT", c.ca, c.cb, c.va, c.vb)
    lvl = {"CCSD": 2, "CCSDT": 3, "CCSDTQ": 3, "CCSDt": 3}[method]
    for k in t:
        if osolver.BLOCKS[k][0] <= lvl:
            t[k] = antisym_random(rng, k, c.ca, c.cb, c.va, c.vb, 0.03)
    kw = dict(shift=0.03)
    if method == "CCSDt":
        from oracle import active as A
        kw.update(act_occ=1, act_unocc=2)
        act = A.ActiveSpace(0, c.oa, c.ob, c.n, 1, 2)
        for k in ("t3aaa", "t3aab", "t3abb", "t3bbb"):
            t[k][~A.p_mask(k, t[k].shape, act)] = 0.0
        vec = A.tvec(t, act, c.ca, c.cb, c.va, c.vb)
    else:
        if method == "CCSD":
            for k in ("t3aaa", "t3aab", "t3abb", "t3bbb"):
                t[k] = np.zeros_like(t[k])
        vec = E.join_tvec(t, c.ca, c.cb, c.va, c.vb)
    out = []
    for pack in ("0", "1"):
        eng = _engine_env(c, method, monkeypatch, {"CCQ_HIP_PACK": pack, "CCQ_HIP_GRAPH": "0"}, **kw)
        eng.set_t(vec)
        eng.profile(True)
        eng.update()
        rep = eng.profile_report()
        eng.profile(False)
        packed = [k for k in rep if k.startswith("gemm:") and "p" in k.rsplit("/", 1)[1]]
        assert bool(packed) == (pack == "1"), sorted(rep)[:12]
        es = [eng.energy()]
        for _ in range(2):
            eng.update()
            es.append(eng.energy())
        out.append((np.array(es), eng.get_t(), eng.get_t4() if method == "CCSDTQ" else None))
        eng.close()
    (e0, t0, q0), (e1, t1, q1) = out
    np.testing.assert_allclose(e1, e0, rtol=1e-11, atol=0)
    assert np.abs(t1 - t0).max() <= 1e-11 * np.abs(t0).max()
    if q0 is not None:
        assert np.abs(q1 - q0).max() <= 1e-11 * max(np.abs(q0).max(), 1e-12)


@pytest.mark.parametrize("method,n,o,ob", [("CCSD", 64, 8, 8), ("CCSD", 64, 8, 7)])
def test_chained_launches_match_single(method, n, o, ob, monkeypatch):
    """GEMM chains (consecutive products into one output with one tile
    mapping, run over their concatenated K) against one launch per GEMM
    (CCQ_HIP_CHAIN=0): three sweeps agree to 1e-11, graph replays of the
    chained sweep equal its eager launches."""
    _say(f"case chained launches {method} n={n} o={o} ob={ob}")
    c = Case(n, o, 2 * n, 41, occ_b=ob)
    vec = _state(c, method, 7)
    out = []
    for chain in ("0", "1"):
        eng = _engine_env(c, method, monkeypatch, {"CCQ_HIP_CHAIN": chain, "CCQ_HIP_GRAPH": "0"}, shift=0.02)
        eng.set_t(vec)
        eng.profile(True)
        eng.update()
        rep = eng.profile_report()
        eng.profile(False)
        chained = [k for k in rep if k.startswith("gemm:") and "c" in k.rsplit("/", 1)[1]]
        if chain == "1":
            _say("  chained classes: " + ", ".join(chained[:6]))
        assert bool(chained) == (chain == "1"), sorted(rep)[:12]
        es = [eng.energy()]
        for _ in range(2):
            eng.update()
            es.append(eng.energy())
        out.append((np.array(es), eng.get_t()))
        eng.close()
    (e0, t0), (e1, t1) = out
    np.testing.assert_allclose(e1, e0, rtol=1e-11, atol=0)
    assert np.abs(t1 - t0).max() <= 1e-11 * np.abs(t0).max()
    engines = [_engine_env(c, method, monkeypatch, {"CCQ_HIP_GRAPH": g}, shift=0.02) for g in ("1", "0")]
    for eng in engines:
        eng.set_t(vec)
    for rep in range(3):
        res = []
        for eng in engines:
            eng.update()
            res.append((eng.energy(), eng.get_t()))
        assert res[0][0] == res[1][0]
        np.testing.assert_array_equal(res[0][1], res[1][1])
    for eng in engines:
        eng.close()


@pytest.mark.parametrize("switch", ["CCQ_HIP_FUSE", "CCQ_HIP_VIEWPERM"])
@pytest.mark.parametrize("method,n,o,ob", [("CCSD", 24, 5, 4), ("CCSDT", 12, 3, 2)])
def test_switch_off_matches_default(switch, method, n, o, ob, monkeypatch):
    """Add folding and the consumer-ordered integral views change only the
    launch structure and memory order: with the switch off, three sweeps
    agree with the default to 1e-11."""
    _say(f"case switch {switch} {method} n={n} o={o} ob={ob}")
    c = Case(n, o, 2 * n, 43, occ_b=ob)
    vec = _state(c, method, 13)
    out = []
    for val in ("0", "1"):
        eng = _engine_env(c, method, monkeypatch, {switch: val, "CCQ_HIP_GRAPH": "0"}, shift=0.02)
        eng.set_t(vec)
        es = []
        for _ in range(3):
            eng.update()
            es.append(eng.energy())
        out.append((np.array(es), eng.get_t()))
        eng.close()
    (e0, t0), (e1, t1) = out
    np.testing.assert_allclose(e1, e0, rtol=1e-11, atol=0)
    assert np.abs(t1 - t0).max() <= 1e-11 * np.abs(t0).max()


@pytest.mark.parametrize("prec", [1, 2])
def test_fp32_grouped_and_chained_close_to_fp64(prec, monkeypatch):
    """The FP32 forms of the grouped and chained launches (mixed precision,
    FP32 contractions on FP64 or FP32 storage): one FP32 sweep at a size
    where both forms run stays within 1e-5 (relative to the step) of the
    FP64 sweep and differs from it."""
    _say(f"case fp32 grouped and chained, precision {prec}")
    c = Case(64, 8, 128, 47)
    vec = _state(c, "CCSD", 17)
    out = []
    for p in (0, prec):
        eng = _engine_env(c, "CCSD", monkeypatch, {"CCQ_HIP_GRAPH": "0"}, shift=0.02)
        eng.set_t(vec)
        eng.set_precision(p)
        eng.profile(True)
        eng.update()
        rep = eng.profile_report()
        eng.profile(False)
        tags = [k.rsplit("/", 1)[1] for k in rep if k.startswith("gemm:")]
        assert any(t.endswith("g") for t in tags) and any("c" in t for t in tags), sorted(rep)[:12]
        out.append(eng.get_t())
        eng.close()
    step = np.abs(out[0] - vec).max()
    diff = np.abs(out[1] - out[0]).max()
    assert diff > 1e-13, "the FP32 contractions did not run"
    assert diff < 1e-5 * step, (diff, step)


@pytest.mark.parametrize("method,n,o,ob", [("CCSD", 24, 5, 4), ("CCSDT", 12, 3, 2),
                                           ("CCSDTQ", 8, 3, 2), ("CCSDt", 10, 3, 2)])
def test_concurrent_levels_match_one_stream(method, n, o, ob, monkeypatch):
    """The launches of a dependency level run on several streams (as a
    fork / join inside the captured sweep graph); they are independent, so
    three sweeps equal the one-stream order (CCQ_HIP_STREAMS=1) bit for bit,
    eager and as graph replays."""
    _say(f"case concurrent levels {method} n={n} o={o} ob={ob}")
    c = Case(n, o, 2 * n, 53, occ_b=ob)
    rng = np.random.default_rng(19)
    t = osolver.zero_amplitudes("CCSDT", c.ca, c.cb, c.va, c.vb)
    lvl = {"CCSD": 2, "CCSDT": 3, "CCSDTQ": 3, "CCSDt": 3}[method]
    for k in t:
        if osolver.BLOCKS[k][0] <= lvl:
            t[k] = antisym_random(rng, k, c.ca, c.cb, c.va, c.vb, 0.03)
    kw = dict(shift=0.03)
    if method == "CCSDt":
        from oracle import active as A
        kw.update(act_occ=1, act_unocc=2)
        act = A.ActiveSpace(0, c.oa, c.ob, c.n, 1, 2)
        for k in ("t3aaa", "t3aab", "t3abb", "t3bbb"):
            t[k][~A.p_mask(k, t[k].shape, act)] = 0.0
        vec = A.tvec(t, act, c.ca, c.cb, c.va, c.vb)
    else:
        if method == "CCSD":
            for k in ("t3aaa", "t3aab", "t3abb", "t3bbb"):
                t[k] = np.zeros_like(t[k])
        vec = E.join_tvec(t, c.ca, c.cb, c.va, c.vb)
    runs = [("1", "0"), ("4", "0"), ("4", "1")]
    engines = [_engine_env(c, method, monkeypatch, {"CCQ_HIP_STREAMS": s, "CCQ_HIP_GRAPH": g}, **kw)
               for s, g in runs]
    for eng in engines:
        eng.set_t(vec)
    for rep in range(3):
        res = []
        for eng in engines:
            eng.update()
            res.append((eng.energy(), eng.get_t(), eng.get_t4() if method == "CCSDTQ" else None))
        for r in res[1:]:
            assert r[0] == res[0][0], rep
            np.testing.assert_array_equal(r[1], res[0][1])
            if method == "CCSDTQ":
                np.testing.assert_array_equal(r[2], res[0][2])
    for eng in engines:
        eng.close()


@pytest.mark.parametrize("n,o,ob,ao,au", [(10, 3, 2, 1, 2), (12, 4, 4, 2, 3), (12, 3, 3, 3, 9)])
def test_pspace_products_match_full_blocks(n, o, ob, ao, au, monkeypatch):
    """CCSDt: the triples residual products computed only on the tuples the
    masked update reads (CCQ_HIP_PSPACE=1, default) against full residual
    blocks (CCQ_HIP_PSPACE=0): three sweeps agree to 1e-11, with fewer
    flops (the same when every orbital is active)."""
    _say(f"case P-space products n={n} o={o} ob={ob} act {ao}/{au}")
    from oracle import active as A
    c = Case(n, o, 2 * n, 59, occ_b=ob)
    rng = np.random.default_rng(23)
    t = osolver.zero_amplitudes("CCSDT", c.ca, c.cb, c.va, c.vb)
    for k in t:
        if osolver.BLOCKS[k][0] <= 3:
            t[k] = antisym_random(rng, k, c.ca, c.cb, c.va, c.vb, 0.03)
    act = A.ActiveSpace(0, c.oa, c.ob, c.n, ao, au)
    for k in ("t3aaa", "t3aab", "t3abb", "t3bbb"):
        t[k][~A.p_mask(k, t[k].shape, act)] = 0.0
    vec = A.tvec(t, act, c.ca, c.cb, c.va, c.vb)
    out = []
    for ps in ("0", "1"):
        eng = _engine_env(c, "CCSDt", monkeypatch, {"CCQ_HIP_PSPACE": ps, "CCQ_HIP_GRAPH": "0"},
                          shift=0.03, act_occ=ao, act_unocc=au)
        eng.set_t(vec)
        eng.profile(True)
        eng.update()
        flops = sum(v["flops"] for v in eng.profile_report().values())
        eng.profile(False)
        es = [eng.energy()]
        for _ in range(2):
            eng.update()
            es.append(eng.energy())
        out.append((np.array(es), eng.get_t(), flops))
        eng.close()
    (e0, t0, f0), (e1, t1, f1) = out
    _say(f"  sweep flops: full {f0 / 1e6:.1f} M, P space {f1 / 1e6:.1f} M")
    np.testing.assert_allclose(e1, e0, rtol=1e-11, atol=0)
    assert np.abs(t1 - t0).max() <= 1e-11 * np.abs(t0).max()
    if ao < o or au < c.va:
        assert f1 < f0


@pytest.mark.parametrize("method,n,o,ob", [("CCSD", 24, 5, 4), ("CCSDT", 12, 3, 2)])
def test_autotuned_plans_match_model_plans(method, n, o, ob, monkeypatch):
    """The measuring autotuner (CCQ_HIP_TUNE=1) only changes tile shapes and
    split-K of the products: three sweeps agree with the model's plans to
    1e-11, and a second context of the process reuses the first one's
    choices (bit-identical sweeps)."""
    _say(f"case autotuned plans {method} n={n} o={o} ob={ob}")
    c = Case(n, o, 2 * n, 61, occ_b=ob)
    vec = _state(c, method, 29)
    out = []
    for tune in ("0", "1", "1"):
        eng = _engine_env(c, method, monkeypatch, {"CCQ_HIP_TUNE": tune, "CCQ_HIP_GRAPH": "0"}, shift=0.02)
        eng.set_t(vec)
        es = []
        for _ in range(3):
            eng.update()
            es.append(eng.energy())
        out.append((np.array(es), eng.get_t()))
        eng.close()
    (e0, t0), (e1, t1), (e2, t2) = out
    np.testing.assert_allclose(e1, e0, rtol=1e-11, atol=0)
    assert np.abs(t1 - t0).max() <= 1e-11 * np.abs(t0).max()
    np.testing.assert_array_equal(e2, e1)
    np.testing.assert_array_equal(t2, t1)


@pytest.mark.parametrize("method,n,o,ob,rhf", [("CCSD", 24, 5, 5, True), ("CCSDT", 12, 3, 3, True),
                                               ("CCSDT", 12, 3, 2, False)])
def test_fp32_storage_autotuned_plans(method, n, o, ob, rhf, monkeypatch):
    """With the autotuner on, the sweep's products also get plans measured in
    FP32 storage (precision mode 2): three mode-2 sweeps with those plans
    agree with mode-2 sweeps on the FP64 plans (CCQ_HIP_TUNE32=0) to FP32
    rounding (1e-6 of the largest amplitude) and stay within 1e-5 of the
    FP64 sweeps' change of the amplitudes."""
    _say(f"case fp32-storage tuned plans {method} n={n} o={o} ob={ob} rhf={rhf}")
    c = Case(n, o, 2 * n, 67, occ_b=ob)
    vec = _state(c, method, 31)
    if rhf:
        t = E.split_tvec(vec, c.ca, c.cb, c.va, c.vb)
        t["t1b"] = t["t1a"].copy()
        t["t2bb"] = t["t2aa"].copy()
        if method == "CCSDT":
            t["t3bbb"] = t["t3aaa"].copy()
            t["t3abb"] = osolver.tran3bto3c(t["t3aab"])
        vec = E.join_tvec(t, c.ca, c.cb, c.va, c.vb)
    out = []
    for prec, tune32 in ((0, "1"), (2, "1"), (2, "0")):
        eng = _engine_env(c, method, monkeypatch, {"CCQ_HIP_TUNE": "1", "CCQ_HIP_TUNE32": tune32,
                                                   "CCQ_HIP_GRAPH": "1"}, shift=0.02, rhf=rhf)
        eng.set_t(vec)
        eng.set_precision(prec)
        for _ in range(3):
            eng.update()
        out.append(eng.get_t())
        eng.close()
    t64, t32, t32m = out
    scale = np.abs(t64).max()
    step = np.abs(t64 - vec).max()
    assert np.abs(t32 - t32m).max() <= 1e-6 * scale
    assert 1e-13 < np.abs(t32 - t64).max() < 1e-5 * step
