"""Seeded synthetic molecular-integral systems in ccq's onebody/twobody form.

ccq reads a one-electron file of ``n(n+1)/2`` lines ``value index`` in the order
i = 1..n, j = 1..i (src/integrals.f90:33-39) and a two-electron file of lines
``i a j b value`` in chemists' order (ia|jb), stored verbatim into
``e2int(i, j, a, b)`` = <ij|ab> with no symmetrisation, closed by the line
``0 0 0 0 E_nuc`` (src/integrals.f90:76-86).  Because ccq never symmetrises,
the writer lists every non-zero index permutation.

Construction (all in FP64):

* occupied orbital energies ~ U[-1.5, -0.4], virtual ones ~ U[0.2, 2.5],
  each sorted ascending;
* B^L_pq = B^L_qp ~ N(0, 0.05^2) for L = 1..nL; (pq|rs) = sum_L B^L_pq B^L_rs,
  which has 8-fold permutational symmetry and is positive semi-definite;
* h_pq = eps_p delta_pq - sum_{k<o} [2 (pq|kk) - (pk|kq)], so the closed-shell
  Fock operator ccq builds (src/integrals.f90:300-364) is diag(eps) up to
  round-off;
* optional irreps (point groups with 8 one-dimensional irreps, D2h): each
  orbital gets an irrep in 0..7 drawn uniformly; every (pq|rs) with
  g_p ^ g_q ^ g_r ^ g_s != 0 is zeroed (the irreps of D2h multiply as XOR
  of their 3-bit labels).

The random stream is numpy's PCG64 ``default_rng(seed)`` drawn in the order
occupied energies, virtual energies, B tensor, irreps.
"""

from __future__ import annotations

import dataclasses
import io
import os

import numpy as np


@dataclasses.dataclass
class System:
    """A closed-shell system in spatial orbitals (all orbitals correlated)."""

    norb: int
    nocc: int
    h: np.ndarray        # (n, n) one-electron integrals
    eri: np.ndarray      # (n, n, n, n) chemists' (pq|rs)
    enuc: float
    eps: np.ndarray      # orbital energies used by the construction
    irreps: np.ndarray | None = None   # (n,) irrep labels 0..7, or None

    @property
    def nvir(self) -> int:
        return self.norb - self.nocc

    def e2int(self) -> np.ndarray:
        """ccq's e2int(i,j,a,b) = <ij|ab> = (ia|jb) as a C-order array [i,j,a,b]."""
        return np.ascontiguousarray(self.eri.transpose(0, 2, 1, 3))


def make_system(norb: int, nocc: int, nL: int, seed: int,
                d2h_irreps: bool = False, enuc: float = 0.0, sigma: float = 0.05) -> System:
    rng = np.random.default_rng(seed)
    nvir = norb - nocc
    eps_o = np.sort(rng.uniform(-1.5, -0.4, nocc))
    eps_v = np.sort(rng.uniform(0.2, 2.5, nvir))
    eps = np.concatenate([eps_o, eps_v])
    raw = rng.normal(0.0, sigma, size=(nL, norb, norb))
    B = np.triu(raw) + np.transpose(np.triu(raw, 1), (0, 2, 1))
    irreps = None
    if d2h_irreps:
        irreps = rng.integers(0, 8, size=norb)
    Bm = B.reshape(nL, norb * norb)
    eri = (Bm.T @ Bm).reshape(norb, norb, norb, norb)
    if irreps is not None:
        # (pq|rs) survives only if g_p^g_q == g_r^g_s; the masked matrix is a
        # direct sum of principal blocks of a PSD matrix, so it stays PSD.
        g = irreps
        pq = g[:, None] ^ g[None, :]
        mask = (pq[:, :, None, None] == pq[None, None, :, :])
        eri = np.where(mask, eri, 0.0)
    occ = slice(0, nocc)
    J = np.einsum('pqkk->pq', eri[:, :, occ, occ])
    K = np.einsum('pkkq->pq', eri[:, occ, occ, :])
    h = np.diag(eps) - (2.0 * J - K)
    h = 0.5 * (h + h.T)
    return System(norb=norb, nocc=nocc, h=h, eri=eri, enuc=float(enuc), eps=eps,
                  irreps=irreps)


def regime_sigma(norb: int, nocc: int, nL: int) -> float:
    """B-entry standard deviation that keeps o^2 v^2 L sigma^4 -- the size of the
    MP2 pair-energy sum -- equal to that of the 24-orbital / 5-occupied / L=48 /
    sigma=0.05 system, whose CC iteration converges (E_corr ~ -0.86 Eh).

    With sigma = 0.05 the 100-orbital / 10-occupied / L=200 system has an MP2
    energy of order -3e2 Eh and the CC iteration diverges within three sweeps
    on the CPU oracle and on the GPU alike, so its time-to-convergence is
    undefined; all other parameters of that workload are kept.
    """
    v = norb - nocc
    ref = 5 ** 2 * 19 ** 2 * 48 * 0.05 ** 4
    return float((ref / (nocc ** 2 * v ** 2 * nL)) ** 0.25)


# The workloads this engine is measured on (spatial orbitals, occupied, L, seed).
# Named synthetic workloads (sizes n/o, auxiliary vectors L, seed, method).
WORKLOADS = {
    "ccsd-n24": dict(norb=24, nocc=5, nL=48, seed=0, method="CCSD"),
    "ccsd-n100": dict(norb=100, nocc=10, nL=200, seed=1, method="CCSD",
                      sigma=regime_sigma(100, 10, 200)),
    "ccsdt-n40": dict(norb=40, nocc=5, nL=80, seed=2, method="CCSDT"),
    "ccsdt-act-n80": dict(norb=80, nocc=7, nL=48, seed=3, method="CCSDt", d2h_irreps=True),
    "ccsdtq-n20": dict(norb=20, nocc=4, nL=40, seed=4, method="CCSDTQ"),
}


def workload_system(name: str) -> System:
    c = dict(WORKLOADS[name])
    c.pop("method")
    return make_system(**c)


def write_onebody(path: str, h: np.ndarray) -> None:
    n = h.shape[0]
    buf = io.StringIO()
    idx = 0
    for i in range(n):
        for j in range(i + 1):
            idx += 1
            buf.write(f"{h[i, j]: .17E} {idx:6d}\n")
    with open(path, "w") as f:
        f.write(buf.getvalue())


def write_twobody(path: str, eri: np.ndarray, enuc: float, thresh: float = 0.0) -> None:
    """Write every (ia|jb) with |value| > thresh as ``i a j b value`` (1-based)."""
    idx = np.argwhere(np.abs(eri) > thresh)
    vals = eri[tuple(idx.T)]
    with open(path, "w") as f:
        block = 1 << 20
        for s in range(0, len(vals), block):
            ii = idx[s:s + block] + 1
            vv = vals[s:s + block]
            lines = [f"{a:5d}{b:5d}{c:5d}{d:5d} {v: .17E}\n"
                     for (a, b, c, d), v in zip(ii.tolist(), vv.tolist())]
            f.write("".join(lines))
        f.write(f"{0:5d}{0:5d}{0:5d}{0:5d} {enuc: .17E}\n")


def write_inputs(dirpath: str, system: System, calc_type: str, **opts) -> str:
    """Write onebody.inp, twobody.inp and a ccq keyword file; return its path."""
    os.makedirs(dirpath, exist_ok=True)
    write_onebody(os.path.join(dirpath, "onebody.inp"), system.h)
    write_twobody(os.path.join(dirpath, "twobody.inp"), system.eri, system.enuc)
    lines = [
        "# System information",
        "froz=0",
        f"nel={2 * system.nocc}",
        f"nvir={2 * system.nvir}",
        f"onebody={os.path.join(dirpath, 'onebody.inp')}",
        f"twobody={os.path.join(dirpath, 'twobody.inp')}",
        "",
        "# Calculation options",
        f"calc_type={calc_type}",
    ]
    for k, v in opts.items():
        lines.append(f"{k}={v}")
    path = os.path.join(dirpath, f"{calc_type}.inp")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path
