"""ccq's Python front door (ccq_amd/ccq_py.py, mirroring tools/ccq_py) on the
GPU: ccq's own example (tools/ccq_py/example_txt_ints.py: H4, CCSDT,
tol 1e-10) reproduces the published H4 CCSDT total (tests/test_small.py:44-47)
from the integral files and from arrays handed over with set_ints."""

import json
import os

import pytest

from oracle import integrals as I
from ccq_amd.ccq_py import Calculation, compute

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "h4_sto3g")


def test_compute_runs_ccq_example(tmp_path):
    print("case ccq_py compute H4 CCSDT", flush=True)
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    compute.configure(0, 4, 4, os.path.join(GOLD, "onebody.inp"),
                      os.path.join(GOLD, "twobody.inp"), False, 1e-10, "CCSDT",
                      log=str(tmp_path / "ccq_py.log"))
    hf, corr, tot = compute.run_calculation()
    assert abs(hf - kat["reference_energy"]) < 1e-9
    assert abs(tot - kat["CCSDT_total"]) < 1e-9
    assert abs(hf + corr - tot) < 1e-14
    assert "CCSDT" in (tmp_path / "ccq_py.log").read_text()
    assert compute.t_vec is not None and compute.t_vec.size > 0


def test_set_ints_and_calculation_class(tmp_path):
    print("case ccq_py set_ints + Calculation H4 CCSDTQ", flush=True)
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    e1 = I.load_e1int(os.path.join(GOLD, "onebody.inp"), 4)
    e2, enuc = I.load_e2int(os.path.join(GOLD, "twobody.inp"), 4)
    compute.set_ints(4, e1, e2, enuc)
    calc = Calculation(nel=4, nvirt=4, calc_type="CCSDTQ", conv_tol=1e-11, rhf=True,
                       log=str(tmp_path / "q.log"))
    ref, corr, tot = calc.run()
    assert abs(tot - kat["CCSDTQ_total"]) < 1e-9
    assert calc.total_en == tot
