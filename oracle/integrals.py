"""Restatement of ccq's integral input, Fock build and HF energy (oracle).

Follows src/integrals.f90 and src/energy.f90 of piecuch-group/ccq.  Arrays
are C-order numpy arrays indexed exactly like ccq's Fortran arrays
(``e2int[i, j, a, b]`` here is ccq's ``e2int(i+1, j+1, a+1, b+1)``).
"""

from __future__ import annotations

import numpy as np


def load_e1int(path: str, norb: int) -> np.ndarray:
    """src/integrals.f90:11-43 -- n(n+1)/2 lines ``value index``, i=1..n, j=1..i."""
    vals = []
    with open(path) as f:
        for line in f:
            if line.strip():
                vals.append(float(line.split()[0].replace("D", "E").replace("d", "e")))
    e1 = np.zeros((norb, norb))
    k = 0
    for i in range(norb):
        for j in range(i + 1):
            e1[i, j] = vals[k]
            e1[j, i] = vals[k]
            k += 1
    return e1


def count_onebody_orbitals(path: str) -> int:
    """src/integrals.f90:246-249 -- orbital count from the number of lines."""
    with open(path) as f:
        nlines = sum(1 for _ in f)
    return int(-(1 - np.sqrt(float(1 + 8 * nlines))) / 2)


def load_e2int(path: str, norb: int) -> tuple[np.ndarray, float]:
    """src/integrals.f90:45-91 -- lines ``i a j b value``; e2int(i,j,a,b) = value.

    No symmetrisation is applied; the line with i+a+j+b == 0 carries the
    nuclear repulsion and ends the read.
    """
    e2 = np.zeros((norb, norb, norb, norb))
    enuc = 0.0
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) < 5:
                break
            i, a, j, b = (int(x) for x in parts[:4])
            hh = float(parts[4].replace("D", "E").replace("d", "e"))
            if i + a + j + b == 0:
                enuc = hh
                break
            e2[i - 1, j - 1, a - 1, b - 1] = hh
    return e2, enuc


def _fortran_real(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def fcidump_header(path: str):
    """src/parser.f90:258-311 (get_fcidump) with parse_key_val
    (src/utils/utils.F90:416-462): NORB and NELEC before the first line
    holding END."""
    norb = nelec = -1
    with open(path) as f:
        for line in f:
            if "END" in line:
                break
            for key in ("NORB", "NELEC"):
                k = line.find(key)
                if k < 0:
                    continue
                rest = line[k:]
                comma = rest.find(",")
                eq = rest.find("=")
                if eq < 0 or (comma >= 0 and eq > comma):
                    continue
                val = rest[eq + 1:comma] if comma >= 0 else rest[eq + 1:]
                if key == "NORB":
                    norb = int(val.split()[0])
                else:
                    nelec = int(val.split()[0])
    return norb, nelec


def load_fcidump(path: str, norb: int):
    """src/integrals.f90:93-210.  Returns (e1int, e2int, E_nuc).

    First pass: orbital energies from lines ``e i 0 0 0``; orbitals are
    renumbered by a stable ascending rank of those energies
    (insertion_rank, src/external/ranking.f90:84-140; energies absent from
    the file count as 0).  Second pass: two-electron lines fill the eight
    permutations of e2int(i,j,a,b) = (ia|jb), one-electron lines
    e1int(i,a) = e1int(a,i), the all-zero line is E_nuc."""
    body = []
    with open(path) as f:
        lines = f.read().splitlines()
    k = 0
    while k < len(lines):
        k += 1
        if "END" in lines[k - 1]:
            break
    for line in lines[k:]:
        parts = line.replace(",", " ").split()
        if len(parts) < 5:
            break
        body.append((_fortran_real(parts[0]),) + tuple(int(x) for x in parts[1:5]))
    eig = np.zeros(norb)
    for v, i, a, j, b in body:
        if i != 0 and a + j + b == 0:
            eig[i - 1] = v
    rank = list(range(norb))
    for i in range(1, norb):          # stable insertion rank
        tmp, j = rank[i], i - 1
        while j >= 0 and not (eig[rank[j]] - eig[tmp] < 0.0):
            rank[j + 1] = rank[j]
            j -= 1
        rank[j + 1] = tmp
    inv = [0] * norb
    for pos, orb in enumerate(rank):
        inv[orb] = pos
    e1 = np.zeros((norb, norb))
    e2 = np.zeros((norb,) * 4)
    enuc = 0.0
    for v, i, a, j, b in body:
        if i and a and j and b:
            i, a, j, b = inv[i - 1], inv[a - 1], inv[j - 1], inv[b - 1]
            for (p, q, r, s) in ((i, j, a, b), (j, i, b, a), (a, b, i, j), (b, a, j, i),
                                 (a, j, i, b), (b, i, j, a), (i, b, a, j), (j, a, b, i)):
                e2[p, q, r, s] = v
        elif i and a and j + b == 0:
            i, a = inv[i - 1], inv[a - 1]
            e1[i, a] = v
            e1[a, i] = v
        elif i + a + j + b == 0:
            enuc = v
    return e1, e2, enuc


def read_sym(path: str, norb: int):
    """src/symmetry.f90:15-119: point group code (1 C1, 2 Cs, 3 C2v, 4 D2h)
    then one irrep per orbital; a missing file is C1 with all irreps 1."""
    import os
    if not os.path.exists(path):
        return 1, np.ones(norb, dtype=np.int64)
    with open(path) as f:
        toks = [ln.split()[0] for ln in f if ln.split()]
    pg = int(toks[0])
    if pg not in (1, 2, 3, 4):
        raise ValueError("read_sym: ERROR: Point-group symmetry not supported")
    return pg, np.array([int(x) for x in toks[1:norb + 1]], dtype=np.int64)


def antisymmetrize(e2int: np.ndarray, froz: int):
    """src/integrals.f90:280-290 -- v_ab = e2int, v_aa = v_bb = e2int - e2int(i,j,b,a).

    Returned over the correlated range froz..n-1 in each index.
    """
    e = e2int[froz:, froz:, froz:, froz:]
    v_ab = e.copy()
    v_aa = e - e.transpose(0, 1, 3, 2)
    return v_aa, v_ab, v_aa.copy()


def gen_fock(e1int: np.ndarray, e2int: np.ndarray, occ_a: int, occ_b: int):
    """src/integrals.f90:300-364 -- F = Z + G split by spin.

    f_a(i,j) = z(i,j) + sum_{k<=occ_b} <ik|jk> + sum_{k<=occ_a} (<ik|jk> - <ik|kj>)
    f_b(i,j) = z(i,j) + sum_{k<=occ_a} <ik|jk> + sum_{k<=occ_b} (<ik|jk> - <ik|kj>)
    """
    J_a = np.einsum('ikjk->ij', e2int[:, :occ_a, :, :occ_a])
    K_a = np.einsum('ikkj->ij', e2int[:, :occ_a, :occ_a, :])
    J_b = np.einsum('ikjk->ij', e2int[:, :occ_b, :, :occ_b])
    K_b = np.einsum('ikkj->ij', e2int[:, :occ_b, :occ_b, :])
    f_a = e1int + J_b + (J_a - K_a)
    f_b = e1int + J_a + (J_b - K_b)
    # ccq fills the lower triangle and mirrors it (src/integrals.f90:355-356)
    f_a = np.tril(f_a) + np.tril(f_a, -1).T
    f_b = np.tril(f_b) + np.tril(f_b, -1).T
    return f_a, f_b


def hf_energy(e1int: np.ndarray, e2int: np.ndarray, occ_a: int, occ_b: int) -> float:
    """src/energy.f90:49-101 -- reference energy without nuclear repulsion."""
    e = 0.0
    e += sum(e1int[i, i] for i in range(occ_a))
    e += sum(e1int[i, i] for i in range(occ_b))
    for i in range(occ_a):
        for j in range(occ_b):
            e += e2int[i, j, i, j]
    for i in range(occ_a):
        for j in range(occ_a):
            e += 0.5 * (e2int[i, j, i, j] - e2int[i, j, j, i])
    for i in range(occ_b):
        for j in range(occ_b):
            e += 0.5 * (e2int[i, j, i, j] - e2int[i, j, j, i])
    return float(e)


def orbital_energies(e1int: np.ndarray, e2int: np.ndarray, occ_a: int) -> np.ndarray:
    """src/energy.f90:11-47 -- diagonal closed-shell Fock elements."""
    n = e1int.shape[0]
    out = np.empty(n)
    for i in range(n):
        acc = e1int[i, i]
        for j in range(occ_a):
            acc += 2.0 * e2int[i, j, i, j] - e2int[i, j, j, i]
        out[i] = acc
    return out
