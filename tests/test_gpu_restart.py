"""ccq's amplitude files through the C ABI: write_vecs writes t_vec.bin and
the ring slot cc_iter_mat-<k>.bin as Fortran records of the DIIS vector
(src/diis.F90:289-349), read_vec restores it, T4 goes to and from
<bin>_4{a..e}.bin (src/utils/cc_utils.f90:89-145), and a solve resumed from
the files ends on the same energy as an uninterrupted one."""

import os

import numpy as np
import pytest

from ccq_amd import engine as E
from ccq_amd import solve as SV
from gpu_helpers import Case

pytestmark = pytest.mark.gpu


def _record(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], np.int32)[0])
    assert int(np.frombuffer(raw[-4:], np.int32)[0]) == n and len(raw) == n + 8
    return np.frombuffer(raw[4:-4], np.float64)


def _engine(c, method, **kw):
    eng = E.Engine(0, c.oa, c.ob, c.n, method, **kw)
    eng.set_ints(c.fa, c.fb, c.vaa, c.vab, c.vbb)
    return eng


@pytest.mark.parametrize("method", ["CCSD", "CCSDT"])
def test_write_vecs_read_vec_round_trip(tmp_path, method):
    print(f"case restart files {method}", flush=True)
    c = Case(10, 3, 20, 5)
    eng = _engine(c, method)
    SV.jacobi_iter(eng, 1e-12, 8, 5, out=None)
    eng.write_vecs(8, 5, str(tmp_path))
    t = eng.get_t()
    vs = eng.lib.ccq_hip_vec_size(eng.h)
    assert vs == (t.size if method != "CCSD" else 2 * 3 * 7 + 3 * 9 * 49)
    # slot mod(8, 6) = 2
    assert sorted(os.listdir(tmp_path)) == ["cc_iter_mat-2.bin", "t_vec.bin"]
    assert np.array_equal(_record(tmp_path / "t_vec.bin"), t[:vs])
    assert np.array_equal(_record(tmp_path / "cc_iter_mat-2.bin"), t[:vs])
    eng.write_vecs(12, 5, str(tmp_path))
    assert (tmp_path / "cc_iter_mat-6.bin").exists()
    other = _engine(c, method)
    other.read_vec(str(tmp_path / "t_vec.bin"))
    assert np.array_equal(other.get_t(), t)
    assert other.energy() == eng.energy()
    with pytest.raises(E.EngineError):
        _engine(Case(9, 3, 18, 5), method).read_vec(str(tmp_path / "t_vec.bin"))
    eng.close()
    other.close()


def test_resumed_solve_ends_on_the_same_energy(tmp_path):
    print("case restart resume CCSDT", flush=True)
    c = Case(9, 3, 18, 8)
    full = _engine(c, "CCSDT")
    r_full = SV.jacobi_iter(full, 1e-10, 60, 5, out=None)
    assert r_full.converged
    first = _engine(c, "CCSDT")
    SV.jacobi_iter(first, 1e-10, 7, 5, out=None)
    first.write_vecs(7, 5, str(tmp_path))
    e_stop = first.energy()
    resumed = _engine(c, "CCSDT")
    e0 = SV.restart(resumed, str(tmp_path / "t_vec.bin"), out=None)
    assert e0 == e_stop
    r = SV.jacobi_iter(resumed, 1e-10, 60, 5, out=None, en_start=e0)
    assert r.converged and abs(r.en_cor - r_full.en_cor) < 1e-9
    for e in (full, first, resumed):
        e.close()


def test_t4_files_round_trip(tmp_path):
    print("case restart T4 files", flush=True)
    c = Case(7, 3, 14, 3)
    eng = _engine(c, "CCSDTQ")
    SV.jacobi_iter(eng, 1e-12, 4, 5, out=None)
    prefix = str(tmp_path / "tvec.bin")
    eng.write_t4(prefix)
    t4 = eng.get_t4()
    assert np.abs(t4).max() > 0
    blocks = E.split_t4(t4, 3, 3, 4, 4)
    for suf, name in zip("abcde", ("t4aaaa", "t4aaab", "t4aabb", "t4abbb", "t4bbbb")):
        assert np.array_equal(_record(prefix + f"_4{suf}.bin"), blocks[name].ravel())
    other = _engine(c, "CCSDTQ")
    other.read_t4(prefix)
    assert np.array_equal(other.get_t4(), t4)
    with pytest.raises(E.EngineError):
        _engine(c, "CCSDT").write_t4(prefix)
    eng.close()
    other.close()
