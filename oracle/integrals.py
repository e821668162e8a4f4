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
