"""ccq's ACC diagram scale factors (src/parser.f90:529-543; default 1.0 at
src/parser.f90:838-842) reach the engine only at their default: both input
front ends (ccq_amd.inputs.parse_config and the Fortran driver, which reads
its keyword file before it touches the GPU) accept 1.0 in ccq's list syntax
and refuse any other value instead of silently computing plain CC."""

import os
import subprocess

import pytest

from ccq_amd import inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRV = os.path.join(ROOT, "fortran", "ccq_drv")

BASE = "nel = 4\nnvir = 4\ncalc_type = CCSD\nonebody = onebody.inp\ntwobody = twobody.inp\n"


def _write(tmp_path, extra):
    p = tmp_path / "input.inp"
    p.write_text(BASE + extra)
    return str(p)


@pytest.mark.parametrize("line", ["t2t2_t2 = 1.0 1.0 1.0 1.0 1.0", "t2t2_t2 = 5*1.0", "t3_t2 = 1d0, 1.0",
                                  "t1t3_t2 = 4*1.0", "t2t2_t3 = 1.0 1.0 1.0", "t2t3_t3 = 5*1.0"])
def test_default_acc_accepted(tmp_path, line):
    cfg = inputs.parse_config(_write(tmp_path, line + "\n"))
    assert cfg.calc_type == "CCSD"


@pytest.mark.parametrize("line", ["t2t2_t2 = 1.0 0.0 0.5 0.5 0.0", "t3_t2 = 1.0 0.5", "t2t3_t3 = 4*1.0 0.0",
                                  "t2t2_t3 = 1.0"])
def test_scaled_acc_refused(tmp_path, line):
    with pytest.raises(ValueError, match="ACC scale factors"):
        inputs.parse_config(_write(tmp_path, line + "\n"))


@pytest.mark.skipif(not os.path.exists(DRV), reason="fortran/ccq_drv not built (make -C fortran)")
def test_fortran_driver_refuses_scaled_acc(tmp_path):
    r = subprocess.run([DRV, _write(tmp_path, "t2t2_t2 = 1.0 0.0 0.5 0.5 0.0\n")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "ACC scale factors t2t2_t2" in r.stderr
