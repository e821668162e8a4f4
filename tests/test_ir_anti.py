"""The antisymmetric position pairs the IR declares per stored block (``anti``
records, which let the executor compute only the canonical half of a pair
and write the mirrored elements with the opposite sign) hold numerically:
every block of a compiled program that carries pairs -- integral views,
amplitudes, intermediates -- is antisymmetric in them after a sweep of the
numpy evaluator on a random system; residual accumulators carry none."""

import numpy as np
import pytest

from ccq_amd.gen import cc_program, ccsd_program, emit, interp
from gpu_helpers import Case, antisym_random
from oracle import solver as osolver


def _pairs_hold(arr, pairs, what):
    for p, q in pairs:
        axes = list(range(arr.ndim))
        axes[p], axes[q] = q, p
        # relative to the block, with an absolute floor: the evaluator's
        # update leaves ~1e-24 roundoff on tuples with a repeated index
        scale = max(np.abs(arr).max(), 1e-3)
        assert np.abs(arr + arr.transpose(axes)).max() <= 1e-13 * scale, (what, p, q)


@pytest.mark.parametrize("method,n,oa,ob", [("CCSD", 7, 3, 2), ("CCSDT", 6, 3, 2)])
def test_declared_pairs_are_antisymmetric(method, n, oa, ob):
    c = Case(n, oa, 16, 70 + n, occ_b=ob)
    rng = np.random.default_rng(3)
    t = osolver.zero_amplitudes("CCSDT", c.ca, c.cb, c.va, c.vb)
    lvl = 2 if method == "CCSD" else 3
    for k in t:
        if osolver.BLOCKS[k][0] <= lvl:
            t[k] = antisym_random(rng, k, c.ca, c.cb, c.va, c.vb, 0.05)
    if method == "CCSD":
        prog, secs = ccsd_program.build(rhf=False)
    else:
        prog, secs = cc_program.build(method, False)
    dims = interp.Dims(c.ca, c.cb, c.va, c.vb)
    base = dict(Fa=c.fa, Fb=c.fb, Vaa=c.vaa, Vab=c.vab, Vbb=c.vbb)
    amps = {k: v.copy() for k, v in t.items()}
    it = interp.Interpreter(prog, secs, dims, base, amps, shift=0.1)
    checked = 0
    # check every intermediate when its section ends (later sections may
    # reassign it)
    for sec in secs:
        it.run_section(sec)
        if sec.kind != "compute":
            continue
        for op in sec.ops:
            key = op.out
            decl = prog.decls[key.name]
            pairs = emit.anti_pairs(decl, key)
            if decl.kind == "resid":
                continue
            if pairs and key in it.store:
                _pairs_hold(it.store[key], pairs, key.tag())
                checked += 1
    # inputs: integral views and amplitude blocks
    for sec in secs:
        for op in getattr(sec, "ops", []):
            for key, _ in op.ins:
                decl = prog.decls[key.name]
                pairs = emit.anti_pairs(decl, key)
                if pairs and decl.kind in ("base", "amp"):
                    _pairs_hold(np.asarray(it.get(key)), pairs, key.tag())
                    checked += 1
    assert checked > 10


def test_residual_blocks_carry_no_pairs():
    prog, secs = ccsd_program.build(rhf=False)
    text = emit.emit(prog, secs, "CCSD", False)
    resid = {ln.split()[1] for ln in text.splitlines() if ln.startswith("tensor ") and " resid " in ln}
    anti = {ln.split()[1]: ln.split()[2:] for ln in text.splitlines() if ln.startswith("anti ")}
    assert resid and not (resid & set(anti))
    assert anti["V.vvvv.aaaa"] == ["0,1", "2,3"]
    assert anti["tau.oovv.aaaa"] == ["0,1", "2,3"]
    assert "V.vvvv.abab" not in anti and "T2.oovv.abab" not in anti
