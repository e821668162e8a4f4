"""The library's integral readers (ccq_amd/csrc/ints_io.cpp) against the
oracle's line-by-line restatements of ccq's readers (oracle/integrals.py,
src/integrals.f90:11-210, src/symmetry.f90:15-119).  Host code only: no GPU.

Bar: bit-exact arrays (both parse decimal text correctly rounded), the same
nuclear repulsion, and ccq's file semantics -- the "0 0 0 0 E" line ends the
two-electron read, a repeated index keeps its last value, Fortran 'D'
exponents and commas are accepted, FCIDUMP orbitals are renumbered by their
energies, a missing symmetry file is C1.
"""

import os

import numpy as np
import pytest

from oracle import integrals as I
from ccq_amd import inputs
from ccq_amd import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden", "h4_sto3g")


def test_h4_golden_files():
    n = inputs.onebody_orbitals(os.path.join(GOLD, "onebody.inp"))
    assert n == I.count_onebody_orbitals(os.path.join(GOLD, "onebody.inp")) == 4
    e1 = inputs.read_onebody(os.path.join(GOLD, "onebody.inp"), n)
    assert np.array_equal(e1, I.load_e1int(os.path.join(GOLD, "onebody.inp"), n))
    e2, enuc = inputs.read_twobody(os.path.join(GOLD, "twobody.inp"), n)
    w2, wnuc = I.load_e2int(os.path.join(GOLD, "twobody.inp"), n)
    assert np.array_equal(e2, w2) and enuc == wnuc


@pytest.mark.parametrize("norb,nocc", [(12, 3), (34, 5)])
def test_synthetic_files_round_trip(tmp_path, norb, nocc):
    """34 orbitals: 1.3 M lines, 60 MB, several parser threads."""
    s = synth.make_system(norb, nocc, 2 * norb, 17, enuc=3.25)
    synth.write_onebody(str(tmp_path / "onebody.inp"), s.h)
    synth.write_twobody(str(tmp_path / "twobody.inp"), s.eri, s.enuc)
    e1 = inputs.read_onebody(str(tmp_path / "onebody.inp"), norb)
    e2, enuc = inputs.read_twobody(str(tmp_path / "twobody.inp"), norb)
    # 17 significant digits round-trip exactly
    assert np.array_equal(e1, s.h)
    assert np.array_equal(e2, s.e2int())
    assert enuc == 3.25
    if norb <= 12:
        w2, _ = I.load_e2int(str(tmp_path / "twobody.inp"), norb)
        assert np.array_equal(e2, w2)


def test_twobody_file_semantics(tmp_path):
    n = 3
    lines = [
        "1 1 1 1  0.5D+00",
        "1, 2, 1, 2, -2.5d-1",          # commas and a Fortran 'd' exponent
        "2 1 2 1 0.125",
        "1 1 1 1 0.75",                 # repeated index: the last value wins
        "0 0 0 0 1.5",                  # nuclear repulsion, end of the read
        "3 3 3 3 9.0",                  # never read
    ]
    p = tmp_path / "twobody.inp"
    p.write_text("\n".join(lines) + "\n")
    e2, enuc = inputs.read_twobody(str(p), n)
    assert enuc == 1.5
    want = np.zeros((n,) * 4)
    want[0, 0, 0, 0] = 0.75
    want[0, 0, 1, 1] = -0.25            # line "i a j b" -> e2int(i, j, a, b)
    want[1, 1, 0, 0] = 0.125
    assert np.array_equal(e2, want)


def test_twobody_without_repulsion_line_stops_at_unreadable_line(tmp_path):
    p = tmp_path / "twobody.inp"
    p.write_text("1 1 1 1 1.0\n\n2 2 2 2 2.0\n")
    e2, enuc = inputs.read_twobody(str(p), 2)
    assert enuc == 0.0
    assert e2[0, 0, 0, 0] == 1.0 and e2[1, 1, 1, 1] == 0.0


def _write_fcidump(path, s, perm, norb, nelec):
    """An FCIDUMP of system ``s`` with its orbitals listed in the order
    ``perm`` (file orbital k is system orbital perm[k]); unique 8-fold
    integrals, one-electron integrals, orbital energies, E_nuc."""
    out = [f" &FCI NORB={norb},NELEC={nelec},MS2=0,", "  ORBSYM=" + "1," * norb,
           "  ISYM=1,", " &END"]
    eri, h = s.eri, s.h
    for p in range(norb):
        for q in range(p + 1):
            for r in range(norb):
                for t in range(r + 1):
                    if p * (p + 1) // 2 + q < r * (r + 1) // 2 + t:
                        continue
                    v = eri[perm[p], perm[q], perm[r], perm[t]]
                    out.append(f"{v: .17E} {p + 1:4d} {q + 1:4d} {r + 1:4d} {t + 1:4d}")
    for p in range(norb):
        for q in range(p + 1):
            out.append(f"{h[perm[p], perm[q]]: .17E} {p + 1:4d} {q + 1:4d}    0    0")
    for p in range(norb):
        out.append(f"{s.eps[perm[p]]: .17E} {p + 1:4d}    0    0    0")
    out.append(f"{s.enuc: .17E}    0    0    0    0")
    path.write_text("\n".join(out) + "\n")


def test_fcidump_reorders_by_orbital_energy(tmp_path):
    norb, nocc = 8, 2
    s = synth.make_system(norb, nocc, 16, 23, enuc=0.75)
    perm = np.random.default_rng(1).permutation(norb)
    p = tmp_path / "FCIDUMP"
    _write_fcidump(p, s, perm, norb, 2 * nocc)
    assert inputs.fcidump_header(str(p)) == I.fcidump_header(str(p)) == (norb, 2 * nocc)
    h, e2, enuc = inputs.read_fcidump(str(p), norb)
    wh, we2, wnuc = I.load_fcidump(str(p), norb)
    assert np.array_equal(h, wh) and np.array_equal(e2, we2) and enuc == wnuc == 0.75
    # the energies sort the orbitals back into the system's own order
    assert np.array_equal(h, s.h)
    assert np.array_equal(e2, s.e2int())


def test_config_with_fcidump_overrides_electron_counts(tmp_path):
    norb, nocc = 6, 2
    s = synth.make_system(norb, nocc, 12, 5)
    _write_fcidump(tmp_path / "FCIDUMP", s, np.arange(norb), norb, 2 * nocc)
    (tmp_path / "ccsd.inp").write_text("nel=99\nnvir=3\nfcidump=FCIDUMP\ncalc_type=CCSD\n")
    cfg = inputs.parse_config(str(tmp_path / "ccsd.inp"))
    assert (cfg.nel, cfg.nvirt, cfg.orbs) == (2 * nocc, 2 * norb - 2 * nocc, norb)
    ints = inputs.load(cfg)
    assert np.array_equal(ints.e1int, s.h)


def test_sym_file(tmp_path):
    p = tmp_path / "sym.inp"
    irr = [1, 5, 8, 2, 3, 3]
    p.write_text("4\n" + "\n".join(str(x) for x in irr) + "\n")
    pg, got = inputs.read_sym(str(p), len(irr))
    wpg, want = I.read_sym(str(p), len(irr))
    assert pg == wpg == 4 and inputs.POINT_GROUPS[pg] == "D2H"
    assert list(got) == list(want) == irr
    pg, got = inputs.read_sym(str(tmp_path / "absent"), 3)
    assert pg == 1 and list(got) == [1, 1, 1]
    (tmp_path / "bad.inp").write_text("7\n1\n")
    with pytest.raises(ValueError):
        inputs.read_sym(str(tmp_path / "bad.inp"), 1)
