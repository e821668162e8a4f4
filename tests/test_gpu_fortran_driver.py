"""The Fortran driver (ISO_C_BINDING over the C ABI) reproduces ccq's H4 CCSDT
known answer, read from its printed summary exactly like ccq's own test
harness does (tests/utils.py:37-58: scan the output for the method line)."""

import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "h4_sto3g")
DRV = os.path.join(ROOT, "fortran", "ccq_drv")


def run_ccq(method, input_file):
    out = subprocess.check_output([DRV, input_file], timeout=120).decode().split("\n")
    ref_energy, cc_energy = 0.0, 0.0
    for line in reversed(out):
        if method in line or "CC" in line:
            _, energy = line.split()
            cc_energy = float(energy)
        elif "Reference" in line:
            _, energy = line.split()
            ref_energy = float(energy)
            break
    return ref_energy, cc_energy


@pytest.mark.skipif(not os.path.exists(DRV), reason="fortran/ccq_drv not built (make -C fortran)")
def test_driver_h4_ccsdt(tmp_path):
    print("case fortran driver h4 ccsdt", flush=True)
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    inp = tmp_path / "CCSDT.inp"
    inp.write_text("\n".join([
        "# System information", "froz=0", "nel=4", "nvir=4",
        f"onebody={GOLD}/onebody.inp", f"twobody={GOLD}/twobody.inp", "",
        "# Calculation options", "rhf=.true.", "tol=1.0e-11", "max_iterations=60",
        "calc_type=CCSDT", "keep_bin=.false.", ""]))
    ref, tot = run_ccq("CCSDT", str(inp))
    # the H4 integral files carry 10 decimals: E_ref is 2.3e-10 off the
    # published value, so compare the correlation energy
    assert abs(ref - kat["reference_energy"]) < 5e-10
    assert abs((tot - ref) - (kat["CCSDT_total"] - kat["reference_energy"])) < 1e-10
