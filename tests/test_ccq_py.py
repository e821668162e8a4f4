"""ccq_py front door, host-side behaviour (no GPU): configure's calc_type
macros and sizes (tools/ccq_py/ccq_py/compute.f90:36-79,
src/parser.f90:316-432), set_ints shape checks."""

import numpy as np
import pytest

from ccq_amd import ccq_py


def test_configure_sizes_and_macros():
    c = ccq_py._Compute()
    c.configure(1, 10, 30, "a", "b", True, 1e-8, "CCSDT")
    assert c._sizes() == (5, 5, 20)
    assert c._s.froz == 1 and c._s.rhf and c._s.tol == 1e-8
    c.configure(0, 10, 30, "a", "b", False, 1e-8, "CR-CC(2,3)")
    assert ccq_py.METHODS[c._s.calc_type] == "CC(t;3)"
    assert (c._s.act_occ, c._s.act_unocc) == (0, 0)
    with pytest.raises(ValueError):
        c.configure(0, 10, 30, "a", "b", False, 1e-8, "CADFCIQMC")
    with pytest.raises(TypeError):
        c.configure(0, 10, 30, "a", "b", False, 1e-8, "CCSD", bogus=1)
    with pytest.raises(RuntimeError):
        ccq_py._Compute().run_calculation()


def test_set_ints_checks_shapes_and_keeps_first():
    c = ccq_py._Compute()
    with pytest.raises(ValueError):
        c.set_ints(3, np.zeros((3, 3)), np.zeros((3, 3, 3, 2)), 0.0)
    a = np.eye(2)
    c.set_ints(2, a, np.zeros((2,) * 4), 1.0)
    c.set_ints(2, 2 * a, np.zeros((2,) * 4), 2.0)
    assert np.array_equal(c.onebody, a) and c.en_repul == 2.0
