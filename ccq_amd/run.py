"""Run one synthetic workload through the engine and print ccq's output.

    python -m ccq_amd.run --workload ccsdt-act-n80 --cct3 --act-occ 3 --act-unocc 8

Workloads are ccq_amd.synth.WORKLOADS (seeded integrals in ccq's regime);
the method is the workload's, CC(t;3) adds the H-bar / left-CC / crcc23 tail
to an active-space CCSDt workload.  Prints ccq's iteration tables and
summaries (src/printing.F90) and the wall time of every phase.
"""

from __future__ import annotations

import argparse
import sys
import time

from . import engine as E
from . import guess as GS
from . import solve as SV
from . import synth
from .inputs import build_integrals


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", required=True, choices=sorted(synth.WORKLOADS))
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--max-iter", type=int, default=80)
    ap.add_argument("--diis", type=int, default=5)
    ap.add_argument("--act-occ", type=int, default=-1)
    ap.add_argument("--act-unocc", type=int, default=-1)
    ap.add_argument("--cct3", action="store_true", help="CC(t;3) tail after CCSDt")
    ap.add_argument("--mixed", action="store_true", help="FP32 contractions + FP64 refinement")
    ap.add_argument("--guess", choices=("zero", "mp2", "learned"), default="zero")
    ap.add_argument("--dense", action="store_true",
                    help="ignore the workload's orbital irreps (no point-group blocks)")
    args = ap.parse_args(argv)
    spec = synth.WORKLOADS[args.workload]
    method = spec["method"]
    if args.cct3:
        if method != "CCSDt":
            ap.error("--cct3 needs an active-space CCSDt workload")
        method = "CC(t;3)"
    t0 = time.perf_counter()
    s = synth.workload_system(args.workload)
    o, n = s.nocc, s.norb
    ints = build_integrals(s.h, s.e2int(), s.enuc, 0, o, o)
    print(f"workload {args.workload}: {method}, n={n} o={o} v={n - o} "
          f"(system {time.perf_counter() - t0:.1f}s)", flush=True)
    t0 = time.perf_counter()
    eng = E.Engine(0, o, o, n, method, diis_space=args.diis, act_occ=args.act_occ,
                   act_unocc=args.act_unocc)
    if s.irreps is not None and not args.dense:
        eng.set_irreps(s.irreps + 1)     # ccq's sym_file numbering (1-based)
    eng.set_ints(ints.f_a, ints.f_b, ints.v_aa, ints.v_ab, ints.v_bb)
    print(f"engine created and integrals uploaded in {time.perf_counter() - t0:.1f}s", flush=True)
    if args.guess != "zero":
        eng.guess(args.guess, GS.load_weights() if args.guess == "learned" else None)
    mixed = SV.Mixed() if args.mixed else None
    print("Coupled cluster\n---------------" + SV.iter_header(), flush=True)
    t0 = time.perf_counter()
    r = SV.jacobi_iter(eng, args.tol, args.max_iter, args.diis, out=sys.stdout, mixed=mixed)
    t_cc = time.perf_counter() - t0
    e_ref = ints.e_hf + ints.enuc
    print(f"CC iterations: {r.iterations}, {t_cc:.2f}s ({t_cc / max(r.iterations, 1) * 1e3:.1f} ms each)",
          flush=True)
    if args.cct3:
        t0 = time.perf_counter()
        eng.hbar()
        t_h = time.perf_counter() - t0
        t0 = time.perf_counter()
        left = SV.solve_lcc(eng, r.en_cor, args.tol, args.max_iter, args.diis, out=sys.stdout)
        t_l = time.perf_counter() - t0
        t0 = time.perf_counter()
        e23 = eng.cct3()
        t_c = time.perf_counter() - t0
        print(SV.print_cct3(e_ref, r.en_cor, e23))
        print(f"H-bar {t_h:.2f}s; left-CC {left.iterations} iterations {t_l:.2f}s; "
              f"crcc23 {t_c:.2f}s", flush=True)
        print(SV.summary(e_ref, r.en_cor + e23[3], method))
    else:
        print(SV.summary(e_ref, r.en_cor, method))
    eng.close()
    return 0 if r.converged else 1


if __name__ == "__main__":
    sys.exit(main())
