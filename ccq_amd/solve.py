"""Host-side CC iteration over the HIP engine: ccq's jacobi_iter.

Line-for-line the control flow of src/solve_cc.f90:175-269 -- update,
energy, ring record, residuum, DIIS every diis_space+1 iterations, the
iteration print, the ``cc.comm`` stop file (:271-291), the three-consecutive
|dE| < tol convergence rule (:293-316), NaN exit and the max_iter message --
with every amplitude operation running on the GPU through ``Engine``.
Only scalars (energy, residuum) cross to the host per iteration.
"""

from __future__ import annotations

import dataclasses
import math
import os
import sys
import time

from .engine import Engine


@dataclasses.dataclass
class SolveResult:
    en_cor: float
    iterations: int
    converged: bool
    failed: bool
    history: list      # (iter, e_cor, dE, residuum, seconds, FP32 contractions?)


def fmt_iteration(it, e, de, res, secs):
    """src/printing.F90:345 -- '(2x,i4,3(f15.10),i5,'' min'',f5.1,'' s'')'."""
    nmin = int(secs) // 60
    nsec = secs - nmin * 60.0
    return f"  {it:4d}{e:15.10f}{de:15.10f}{res:15.10f}{nmin:5d} min{nsec:5.1f} s"


def iter_header():
    """src/printing.F90:303-310."""
    return f"\n  {'It.':>4}{'E (Corr)':>15}{'dE':>15}{'Residuum':>15}{'Wall Time':>16}\n  " + "-" * 65


def check_sig(path: str = "cc.comm") -> bool:
    """src/solve_cc.f90:271-291: a file ``cc.comm`` stops the loop and is deleted."""
    if os.path.exists(path):
        os.remove(path)
        return True
    return False


@dataclasses.dataclass
class Mixed:
    """Mixed-precision schedule: sweeps run with FP32 MFMA contractions until
    |dE| < switch_de, then FP64 to the end; with refine_every = k every k-th
    sweep before the switch is FP64 as well.  Convergence (three consecutive
    |dE| < tol) is only counted over FP64 sweeps, so the converged energy and
    amplitudes are those of the FP64 equations.  The default switch lies
    above the FP32-contraction noise of dE (1e-6 - 1e-5 on the 40-orbital
    CCSDT and 100-orbital CCSD systems), where the FP32 phase leaves the FP64
    trajectory's iteration count unchanged.  mode: the precision mode of the
    FP32 sweeps (1: FP32 MFMA on FP64 storage; 2: FP32 storage of every
    tensor of the sweep as well, fewer bytes per sweep)."""
    switch_de: float = 3e-5
    refine_every: int = 0
    mode: int = 1


def restart(eng: Engine, path: str = "t_vec.bin", bin_file: str | None = None,
            out=sys.stdout) -> float:
    """Resume from ccq's files: the DIIS vector from ``path`` (t_vec.bin) and,
    for CCSDTQ, T4 from <bin_file>_4{a..e}.bin; prints ccq's restart line
    (src/solve_cc.f90:56-60) and returns the restart correlation energy,
    which seeds the first energy difference of the next jacobi_iter."""
    eng.read_vec(path)
    if bin_file is not None and eng.t4_size:
        eng.read_t4(bin_file)
    e = eng.energy()
    if out:
        print(f"  {'Restart correlation energy':>27s}  {e:16.10f}\n", file=out)
    return e


def jacobi_iter(eng: Engine, tol: float, max_iter: int, diis_space: int,
                out=sys.stdout, on_iter=None, stop_file: str = "cc.comm",
                mixed: Mixed | None = None, en_start: float = 0.0) -> SolveResult:
    conv = [1.0, 1.0, 1.0]
    # conv%en_cor before iteration 1: ccq leaves it uninitialised
    # (src/solver_types.f90:19) and sets it on restart (src/solve_cc.f90:57)
    en_cor = en_start
    failed = False
    converged = False
    hist = []
    prev = time.perf_counter()
    it = 0
    low = mixed is not None
    for it in range(1, max_iter + 1):
        if mixed is not None:
            fp32 = low and not (mixed.refine_every and it % mixed.refine_every == 0)
            eng.set_precision(mixed.mode if fp32 else 0)
        eng.update()
        e_new = eng.energy()
        de = en_cor - e_new
        conv = [conv[1], conv[2], de]
        en_cor = e_new
        eng.diis_record(it)
        res = eng.residuum(it)
        if it % (diis_space + 1) == 0:
            if out:
                print("      DIIS cycle", file=out)
            eng.diis_extrapolate()
        now = time.perf_counter()
        hist.append((it, e_new, de, res, now - prev, mixed is not None and fp32))
        if out:
            print(fmt_iteration(it, e_new, de, res, now - prev), file=out, flush=True)
        prev = now
        if on_iter is not None:
            on_iter(it, e_new)
        if check_sig(stop_file):
            if out:
                print("Signaled to end", file=out)
            break
        if low and abs(de) < mixed.switch_de:
            # final sweeps in FP64; only they count towards convergence
            low = False
            conv = [1.0, 1.0, 1.0]
            if out:
                print("      FP64 refinement", file=out)
        elif mixed is not None and fp32:
            conv = [1.0, 1.0, 1.0]
        if all(abs(c) < tol for c in conv):
            converged = True
            break
        if math.isnan(e_new):
            if out:
                print("FAILED TO CONVERGE", file=out)
            break
        if it == max_iter:
            if out:
                print("FAILED TO CONVERGE IN", max_iter, " ITERATIONS. KEEPING LAST T VECTOR.", file=out)
            failed = True
    if mixed is not None:
        eng.set_precision(0)
    return SolveResult(en_cor=en_cor, iterations=it, converged=converged, failed=failed, history=hist)


def summary(e_ref: float, e_cor: float, calc_type: str) -> str:
    """src/printing.F90:260-284 -- '(a20,f18.12)' lines."""
    lines = ["", "Calculation Summary (Eh)", "------------------------",
             f"{'Reference':>20}{e_ref:18.12f}",
             f"{'Correlation':>20}{e_cor:18.12f}",
             f"{calc_type:>20}{e_ref + e_cor:18.12f}"]
    return "\n".join(lines)


class _LeftIteration:
    """jacobi_iter's five calls for the left-CC equations: solve_lcc binds
    update_ptr => update_l2, calculate_energy_ptr => calculate_left_energy and
    points the DIIS ring at l_vec (src/solve_cc.f90:101-173)."""

    def __init__(self, eng: Engine):
        self.eng = eng

    def update(self):
        self.eng.left_update()

    def energy(self) -> float:
        return self.eng.left_energy()

    def diis_record(self, it: int):
        self.eng.diis_record(it)

    def residuum(self, it: int) -> float:
        return self.eng.residuum(it)

    def diis_extrapolate(self):
        self.eng.diis_extrapolate()


def solve_lcc(eng: Engine, e_cor: float, tol: float, max_iter: int, diis_space: int,
              out=sys.stdout, stop_file: str = "cc.comm") -> SolveResult:
    """Left-CCSD on the resident H-bar (ccq_hip_hbar first), l_vec from zero."""
    if out:
        print("Left Coupled cluster\n--------------------", file=out)
        print(iter_header(), file=out)
    eng.left_init(e_cor)
    eng.left_mode(True)
    try:
        return jacobi_iter(_LeftIteration(eng), tol, max_iter, diis_space, out=out,
                           stop_file=stop_file)
    finally:
        eng.left_mode(False)


@dataclasses.dataclass
class CCT3Result:
    ccsdt: SolveResult
    left: SolveResult
    e23: tuple          # crcc23's (A, B, C, D) corrections

    @property
    def en_cor(self) -> float:
        return self.ccsdt.en_cor + self.e23[3]


def cct3(eng: Engine, tol: float, max_iter: int, diis_space: int, out=sys.stdout) -> CCT3Result:
    """CC(t;3): CCSDt, H-bar, left-CCSD, crcc23 (src/calc_driver.f90:53-72)."""
    r = jacobi_iter(eng, tol, max_iter, diis_space, out=out)
    eng.hbar()
    left = solve_lcc(eng, r.en_cor, tol, max_iter, diis_space, out=out)
    e23 = eng.cct3()
    return CCT3Result(ccsdt=r, left=left, e23=e23)


def print_cct3(e_ref: float, e_cor: float, e23) -> str:
    """print_cct3 (src/printing.F90:234-258)."""
    rows = [("CCSDt", 0.0), ("CC(t;3),A", e23[0]), ("CC(t;3),B", e23[1]),
            ("CC(t;3),C", e23[2]), ("CC(t;3)", e23[3])]
    lines = ["", f"{'Method':>20}{'Correlation (Eh)':>18}{'Total Energy (Eh)':>18}", " " * 10 + "-" * 46]
    for name, de in rows:
        lines.append(f"{name:>20}{e_cor + de:18.12f}{e_ref + e_cor + de:18.12f}")
    return "\n".join(lines)
