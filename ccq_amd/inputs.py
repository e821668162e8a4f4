"""ccq input: keyword file, onebody/twobody integrals, Fock and HF energy.

Keyword semantics follow src/parser.f90 (get_config :99-170, calc macros
:316-432, system data :434-504, run options :638-732, defaults :779-844):
``key = value`` lines, '#'/'!' lines are comments, unknown keys are ignored,
``nel``/``nvir``/``froz`` count spin orbitals, an integer ``tol`` means
10**tol.  Integral files follow src/integrals.f90:11-91; the Fock build and
antisymmetrisation follow src/integrals.f90:280-364; E_HF src/energy.f90:49-101.
"""

from __future__ import annotations

import dataclasses
import os

import numpy as np


@dataclasses.dataclass
class Config:
    calc_type: str = "CCSD"
    froz_spin: int = 0
    nel: int = -1
    nvirt: int = -1
    mult: int = 0
    tol: float = 1.0e-4
    diis_space: int = 5
    max_iter: int = 60
    shift: float = 0.0
    rhf: bool = False
    restart: bool = False
    onebody: str = "onebody.inp"
    twobody: str = "twobody.inp"
    fcidump: str = ""
    sym_file: str = ""
    act_occ: int = 0
    act_unocc: int = 0
    act_ind_t: int = 0
    num_threads: int = 1
    label: str = ""
    keep_bin: bool = True

    # src/parser.f90:476-504 (process_sys_data)
    @property
    def froz(self) -> int:
        return self.froz_spin // 2

    @property
    def occ_a(self) -> int:
        return self.nel // 2

    @property
    def occ_b(self) -> int:
        return self.occ_a - (self.mult - 1) // 2 if self.mult > 0 else self.occ_a

    @property
    def orbs(self) -> int:
        return (self.nel + self.nvirt) // 2


def _logical(val: str) -> bool:
    v = val.strip().lower().strip('.')
    if v in ("true", "t"):
        return True
    if v in ("false", "f"):
        return False
    raise ValueError(f"not a logical: {val!r}")


# ACC diagram scale factors (src/parser.f90:529-543; REAL*4 lists, default
# 1.0 at src/parser.f90:838-842) and their lengths
ACC_KEYS = {"t2t2_t2": 5, "t3_t2": 2, "t1t3_t2": 4, "t2t2_t3": 3, "t2t3_t3": 5}


def _fortran_reals(val: str):
    """Values of a Fortran list-directed real list ('1.0, 1.0', '5*1.0', 1d0)."""
    out = []
    for tok in val.replace(",", " ").split():
        rep, _, x = tok.rpartition("*")
        out += [float(x.replace("d", "e").replace("D", "E"))] * (int(rep) if rep else 1)
    return out


def parse_config(path: str) -> Config:
    cfg = Config()
    base = os.path.dirname(os.path.abspath(path))
    with open(path) as f:
        lines = f.read().splitlines()
    for line in lines:
        if "#" in line or "!" in line:
            continue
        if not line.strip():
            continue
        if "=" in line:
            key, val = line.split("=", 1)
            key, val = key.strip(), val.strip()
        else:
            key, val = line.strip(), ""
        if key in ("core", "nfroz", "frozen", "froz"):
            cfg.froz_spin = int(val)
        elif key in ("nel", "electrons"):
            cfg.nel = int(val)
        elif key in ("nvir", "nvirt", "virtuals"):
            cfg.nvirt = int(val)
        elif key in ("multiplicity", "mult"):
            cfg.mult = int(val)
        elif key == "rhf":
            cfg.rhf = _logical(val)
        elif key == "shift":
            cfg.shift = float(val.replace("d", "e").replace("D", "E"))
        elif key in ("max_iterations", "max_iter"):
            cfg.max_iter = int(val)
        elif key in ("diis_space", "diis"):
            cfg.diis_space = int(val)
        elif key in ("tolerance", "tol", "itol"):
            try:
                cfg.tol = 10.0 ** int(val)
            except ValueError:
                cfg.tol = float(val.replace("d", "e").replace("D", "E"))
        elif key == "calc_type":
            cfg.calc_type = val
        elif key in ("restart", "rest"):
            cfg.restart = True
        elif key == "onebody":
            cfg.onebody = val
        elif key == "twobody":
            cfg.twobody = val
        elif key == "fcidump":
            cfg.fcidump = val
        elif key in ("symfile", "sym_file"):
            cfg.sym_file = val
        elif key == "act_occ":
            cfg.act_occ = int(val)
        elif key == "act_unocc":
            cfg.act_unocc = int(val)
        elif key == "num_threads":
            cfg.num_threads = int(val)
        elif key == "label":
            cfg.label = val
        elif key == "keep_bin":
            cfg.keep_bin = _logical(val)
        elif key in ACC_KEYS:
            # the engine computes plain CC: ACC scaling only at its default
            vals = _fortran_reals(val)
            if len(vals) < ACC_KEYS[key] or any(v != 1.0 for v in vals[:ACC_KEYS[key]]):
                raise ValueError(f"ACC scale factors {key} = {val} are not supported "
                                 "(only the default 1.0)")
        # any other keyword is ignored, as in ccq
    if cfg.mult == 0:
        cfg.mult = 1
    for attr in ("fcidump", "sym_file"):
        p = getattr(cfg, attr)
        if p and not os.path.isabs(p) and not os.path.exists(p):
            cand = os.path.join(base, p)
            if os.path.exists(cand):
                setattr(cfg, attr, cand)
    if cfg.fcidump:
        # the FCIDUMP header overrides nel / nvir (src/parser.f90:168-169,258-311)
        norb, nelec = fcidump_header(cfg.fcidump)
        cfg.nel = nelec
        cfg.nvirt = 2 * norb - nelec
    for attr in ("onebody", "twobody"):
        p = getattr(cfg, attr)
        if p and not os.path.isabs(p) and not os.path.exists(p):
            cand = os.path.join(base, p)
            if os.path.exists(cand):
                setattr(cfg, attr, cand)
    return cfg


def _reader():
    from .engine import load_library
    return load_library()


def _check(rc, lib, what):
    if rc != 0:
        raise ValueError(f"{what}: " + lib.ccq_hip_read_last_error().decode())


def onebody_orbitals(path: str) -> int:
    """Orbital count from the onebody file's n(n+1)/2 records (src/integrals.f90:248-251)."""
    import ctypes
    lib = _reader()
    n = ctypes.c_int64()
    _check(lib.ccq_hip_onebody_orbitals(path.encode(), ctypes.byref(n)), lib, "onebody")
    return int(n.value)


def read_onebody(path: str, norb: int) -> np.ndarray:
    """load_e1int (src/integrals.f90:11-41) through the library's reader."""
    lib = _reader()
    h = np.zeros((norb, norb), order="F")
    _check(lib.ccq_hip_read_onebody(path.encode(), norb, h.ctypes.data), lib, "load_e1int")
    return np.ascontiguousarray(h)


def read_twobody(path: str, norb: int):
    """load_e2int (src/integrals.f90:43-91): (e2int[i,j,a,b] = value of line
    'i a j b', E_nuc), parsed in parallel by the library's reader."""
    import ctypes
    lib = _reader()
    e2 = np.zeros((norb,) * 4, order="F")
    enuc = ctypes.c_double(0.0)
    _check(lib.ccq_hip_read_twobody(path.encode(), norb, e2.ctypes.data, ctypes.byref(enuc)),
           lib, "load_e2int")
    return np.ascontiguousarray(e2), float(enuc.value)


def fcidump_header(path: str):
    """(NORB, NELEC) of an FCIDUMP header (get_fcidump, src/parser.f90:258-311)."""
    import ctypes
    lib = _reader()
    norb, nelec = ctypes.c_int64(), ctypes.c_int64()
    _check(lib.ccq_hip_fcidump_header(path.encode(), ctypes.byref(norb), ctypes.byref(nelec)),
           lib, "get_fcidump")
    return int(norb.value), int(nelec.value)


def read_fcidump(path: str, norb: int):
    """load_fcidump (src/integrals.f90:93-210): (e1int, e2int[i,j,a,b], E_nuc)."""
    import ctypes
    lib = _reader()
    h = np.zeros((norb, norb), order="F")
    e2 = np.zeros((norb,) * 4, order="F")
    enuc = ctypes.c_double(0.0)
    _check(lib.ccq_hip_read_fcidump(path.encode(), norb, h.ctypes.data, e2.ctypes.data,
                                     ctypes.byref(enuc)), lib, "load_fcidump")
    return np.ascontiguousarray(h), np.ascontiguousarray(e2), float(enuc.value)


def read_sym(path: str, norb: int):
    """read_sym (src/symmetry.f90:15-119): (point group 1..4, irreps[orbs], 1-based)."""
    import ctypes
    lib = _reader()
    pg = ctypes.c_int64()
    irr = np.zeros(norb, dtype=np.int64)
    _check(lib.ccq_hip_read_sym(path.encode(), norb, ctypes.byref(pg), irr.ctypes.data),
           lib, "read_sym")
    return int(pg.value), irr


POINT_GROUPS = {1: "C1", 2: "CS", 3: "C2V", 4: "D2H"}


@dataclasses.dataclass
class Integrals:
    e1int: np.ndarray
    e2int: np.ndarray
    enuc: float
    f_a: np.ndarray
    f_b: np.ndarray
    v_aa: np.ndarray
    v_ab: np.ndarray
    v_bb: np.ndarray
    e_hf: float


def build_integrals(e1int, e2int, enuc, froz, occ_a, occ_b) -> Integrals:
    e = e2int[froz:, froz:, froz:, froz:]
    v_ab = np.ascontiguousarray(e)
    v_aa = np.ascontiguousarray(e - e.transpose(0, 1, 3, 2))
    Ja = np.einsum('ikjk->ij', e2int[:, :occ_a, :, :occ_a])
    Ka = np.einsum('ikkj->ij', e2int[:, :occ_a, :occ_a, :])
    Jb = np.einsum('ikjk->ij', e2int[:, :occ_b, :, :occ_b])
    Kb = np.einsum('ikkj->ij', e2int[:, :occ_b, :occ_b, :])
    f_a = e1int + Jb + Ja - Ka
    f_b = e1int + Ja + Jb - Kb
    f_a = np.tril(f_a) + np.tril(f_a, -1).T
    f_b = np.tril(f_b) + np.tril(f_b, -1).T
    oa, ob = slice(0, occ_a), slice(0, occ_b)
    ehf = np.trace(e1int[oa, oa]) + np.trace(e1int[ob, ob])
    ehf += np.einsum('ijij->', e2int[oa, ob, oa, ob])
    d = e2int - e2int.transpose(0, 1, 3, 2)
    ehf += 0.5 * np.einsum('ijij->', d[oa, oa, oa, oa]) + 0.5 * np.einsum('ijij->', d[ob, ob, ob, ob])
    return Integrals(e1int, e2int, enuc, f_a, f_b, v_aa, v_ab, v_aa, float(ehf))


def load(cfg: Config) -> Integrals:
    """load_ints (src/integrals.f90:212-298): FCIDUMP if given, else the
    onebody / twobody pair (whose record count must give the orbital count)."""
    n = cfg.orbs
    if cfg.fcidump:
        e1, e2, enuc = read_fcidump(cfg.fcidump, n)
    else:
        if onebody_orbitals(cfg.onebody) != n:
            raise ValueError("load_ints: Number of orbitals doesn't match the number of integrals")
        e1 = read_onebody(cfg.onebody, n)
        e2, enuc = read_twobody(cfg.twobody, n)
    return build_integrals(e1, e2, enuc, cfg.froz, cfg.occ_a, cfg.occ_b)
