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


@pytest.mark.skipif(not os.path.exists(DRV), reason="fortran/ccq_drv not built (make -C fortran)")
def test_driver_cct3_matches_oracle(tmp_path):
    """calc_type=CCT3 through the driver: CCSDt, H-bar, left-CCSD, crcc23 and
    ccq's print_cct3 table, against the oracle's energies stored in
    tests/golden/cct3_n7_33.npz (same seeded 7-orbital system)."""
    import numpy as np
    from ccq_amd import synth
    print("case fortran driver cct3 n7", flush=True)
    g = np.load(os.path.join(ROOT, "tests", "golden", "cct3_n7_33.npz"))
    s = synth.make_system(7, 3, 20, 11)
    inp = synth.write_inputs(str(tmp_path), s, "CCT3", act_occ=1, act_unocc=1, tol="1.0e-11",
                             diis_space=5, max_iterations=80)
    out = subprocess.check_output([DRV, inp], timeout=120).decode()
    ref, tot = run_ccq("CC(t;3)", inp)
    ecor, e23 = float(g["ecor"]), g["e23"]
    assert abs((tot - ref) - (ecor + e23[3])) < 1e-9
    rows = {}
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] in ("CCSDt", "CC(t;3),A", "CC(t;3),B", "CC(t;3),C",
                                            "CC(t;3)"):
            rows[parts[0]] = float(parts[1])
    assert abs(rows["CCSDt"] - ecor) < 1e-9
    for name, k in (("CC(t;3),A", 0), ("CC(t;3),B", 1), ("CC(t;3),C", 2), ("CC(t;3)", 3)):
        assert abs(rows[name] - (ecor + e23[k])) < 1e-9, name
    assert "Left Coupled cluster" in out


@pytest.mark.skipif(not os.path.exists(DRV), reason="fortran/ccq_drv not built (make -C fortran)")
def test_driver_sym_file_blocks_equal_dense(tmp_path):
    """A D2h system with ccq's sym_file keyword (point group 4, one 1-based
    irrep per orbital, src/symmetry.f90:84-114): the driver hands the irreps
    to the engine, which runs every product in irrep blocks; the printed
    CCSDT energy equals the run without the file (dense products)."""
    import numpy as np
    from ccq_amd import synth
    print("case fortran driver sym_file ccsdt n10", flush=True)
    s = synth.make_system(10, 3, 20, 17, d2h_irreps=True)
    totals = []
    for sym in (False, True):
        d = tmp_path / ("sym" if sym else "dense")
        opts = dict(rhf=".true.", tol="1.0e-10", max_iterations=80)
        if sym:
            d.mkdir()
            (d / "sym.inp").write_text("4\n" + "\n".join(str(int(g) + 1) for g in s.irreps) + "\n")
            opts["sym_file"] = str(d / "sym.inp")
        inp = synth.write_inputs(str(d), s, "CCSDT", **opts)
        ref, tot = run_ccq("CCSDT", inp)
        totals.append(tot)
    assert np.isfinite(totals[0])
    assert abs(totals[1] - totals[0]) < 1e-9
