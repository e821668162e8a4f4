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
    for attr in ("onebody", "twobody"):
        p = getattr(cfg, attr)
        if p and not os.path.isabs(p) and not os.path.exists(p):
            cand = os.path.join(base, p)
            if os.path.exists(cand):
                setattr(cfg, attr, cand)
    return cfg


def read_onebody(path: str, norb: int) -> np.ndarray:
    vals = np.loadtxt(path, usecols=0, converters={0: lambda s: float(s.replace(b"D", b"E") if isinstance(s, bytes) else s.replace("D", "E"))}, ndmin=1)
    n = norb
    if vals.size < n * (n + 1) // 2:
        raise ValueError("onebody file too short for the orbital count")
    h = np.zeros((n, n))
    il = np.tril_indices(n)
    # file order: i = 1..n, j = 1..i  (row-major lower triangle)
    h[il] = vals[: n * (n + 1) // 2]
    h = h + np.tril(h, -1).T
    return h


def read_twobody(path: str, norb: int):
    """Return (e2int[i,j,a,b] = value of line 'i a j b', E_nuc)."""
    raw = np.loadtxt(path, ndmin=2)
    idx = raw[:, :4].astype(np.int64)
    end = np.nonzero(idx.sum(axis=1) == 0)[0]
    enuc = 0.0
    if end.size:
        enuc = float(raw[end[0], 4])
        raw, idx = raw[: end[0]], idx[: end[0]]
    e2 = np.zeros((norb,) * 4)
    i, a, j, b = (idx[:, k] - 1 for k in range(4))
    e2[i, j, a, b] = raw[:, 4]
    return e2, enuc


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
    n = cfg.orbs
    e1 = read_onebody(cfg.onebody, n)
    e2, enuc = read_twobody(cfg.twobody, n)
    return build_integrals(e1, e2, enuc, cfg.froz, cfg.occ_a, cfg.occ_b)
