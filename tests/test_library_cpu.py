"""CPU-side checks of the product: the C-ABI library loads and exports every
symbol include/ccq_hip.h declares; the compiled block program matches the
oracle when evaluated with numpy (no GPU needed)."""

import ctypes
import os
import re

import numpy as np
import pytest

from oracle import solver as osolver
from ccq_amd import engine as E
from ccq_amd.gen import ccsd_program, emit, interp
from gpu_helpers import Case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "ccq_hip.h")).read()
    return sorted(set(re.findall(r"\b(ccq_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(E.LIB_PATH)
    names = header_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(E.EXPORTS), "binding and header disagree"


def test_embedded_ir_is_current():
    for method in emit.METHODS:
        gh = emit.generator_hash(method)
        for rhf in emit.VARIANTS[method]:
            tag = "rhf" if rhf else "uhf"
            path = os.path.join(ROOT, "ccq_amd", "csrc", f"ir_{emit.STEM[method]}_{tag}.inc")
            with open(path) as f:
                assert f"generator {gh}" in f.readline(), path
    for rhf in (False, True):
        prog, secs = ccsd_program.build(rhf=rhf)
        text = emit.emit(prog, secs, "CCSD", rhf)
        tag = "rhf" if rhf else "uhf"
        inc = open(os.path.join(ROOT, "ccq_amd", "csrc", f"ir_ccsd_{tag}.inc")).read()
        assert text in inc


@pytest.mark.parametrize("n,oa,ob,rhf", [(7, 3, 2, False), (7, 3, 3, True)])
def test_compiled_program_sweep_equals_oracle(n, oa, ob, rhf):
    c = Case(n, oa, 20, 40 + n, occ_b=ob)
    t = c.random_t(9)
    if rhf:
        t['t2bb'] = t['t2aa'].copy()
        t['t1b'] = t['t1a'].copy()
    prog, secs = ccsd_program.build(rhf=rhf)
    dims = interp.Dims(c.ca, c.cb, c.va, c.vb)
    base = dict(Fa=c.fa, Fb=c.fb, Vaa=c.vaa, Vab=c.vab, Vbb=c.vbb)
    amps = {k: v.copy() for k, v in t.items()}
    it = interp.Interpreter(prog, secs, dims, base, amps, shift=0.1)
    it.sweep()
    D = osolver.block_denominators(t, c.fa, c.fb, c.ca, c.cb, 0.1)
    osolver.sweep(t, osolver.sweep_order('CCSD', rhf), c.provider(), D)
    for k in ['t1a', 't1b', 't2aa', 't2ab', 't2bb']:
        np.testing.assert_allclose(amps[k], t[k], atol=1e-13, rtol=0, err_msg=k)
