"""Benchmark: time-to-|dE|<1e-7 Eh of the CC iteration on one MI355X,
mixed precision against plain FP64.

Workload (N=1): closed-shell CCSD, 100 spatial orbitals (10 occupied,
90 virtual), synthetic integrals (pq|rs) = sum_L B B with L = 200 from seed 1,
tol-driven ccq iteration (diis_space 5, shift 0).

Protocol
  1. build the system and upload it (untimed);
  2. W warm-up iterations from T = 0 in each precision, then a tight untimed
     FP64 solve to get the converged energy E_conv and amplitudes (tol 1e-11);
  3. plain leg: reset T = 0 and time exactly K FP64 iterations (update,
     energy, DIIS ring, residuum, DIIS every 6th), synchronising the stream
     around the timed region and after every iteration;
  4. learned leg: the learned T2 guess (ccq_amd/guess_weights.npz; the MP2
     guess if absent), computed inside the timed region, then K FP64
     iterations;
  5. mixed+learned leg: the learned guess, then K iterations with FP32-MFMA
     contractions until |dE| < 1e-6 and FP64 after that (the refinement);
  6. value = the mixed+learned leg's wall time until the first iteration
     with |E - E_conv| < 1e-7; every leg's time, iteration count, final
     energy and max-abs T error against the FP64 solution, and the speedup
     over the plain leg go into config.
  Roofline: the dominant GEMM class (the contraction shape with the largest
  share of the sweep) timed with HIP events on the engine's stream during one
  extra profiled iteration; algorithmic FLOPs 2*M*N*K per launch.
  traffic: HBM bytes per launch of that class from the rocprofv3 FETCH_SIZE /
  WRITE_SIZE runs summarised under profiles/ (copy-calibrated), when measured.
  cpu_baseline: the numpy oracle (same equations, same sweep, OpenBLAS) on a
  bounded sample of the same workload, extrapolated to the same iteration count.

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ccq_amd import engine as E  # noqa: E402
from ccq_amd import guess  # noqa: E402
from ccq_amd import synth  # noqa: E402
from ccq_amd.inputs import build_integrals  # noqa: E402

METRIC = "time-to-|dE|<1e-7 Eh, mixed+learned vs plain FP64 (1 GPU); % FP64 MFMA peak"
PEAK_FP64_MFMA = 78.6e12
PEAK_HBM = 8.0e12
SWITCH_DE = 1e-6     # mixed leg: FP32 contractions until |dE| falls below this


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(name: str):
    s = synth.workload_system(name)
    ints = build_integrals(s.h, s.e2int(), s.enuc, 0, s.nocc, s.nocc)
    return s, ints


def run_iterations(eng, nsteps, diis_space, energies=None, stamps=None, switch_de=None):
    """nsteps iterations of ccq's loop.  switch_de: mixed precision -- FP32
    contractions until |dE| < switch_de, then FP64 (solve.Mixed's schedule)."""
    en_cor = 0.0
    low = switch_de is not None
    if low:
        eng.set_precision(1)
    for it in range(1, nsteps + 1):
        eng.update()
        e = eng.energy()
        if low and abs(en_cor - e) < switch_de:
            low = False
            eng.set_precision(0)
        eng.diis_record(it)
        eng.residuum(it)
        if it % (diis_space + 1) == 0:
            eng.diis_extrapolate()
        if energies is not None:
            energies.append(e)
        if stamps is not None:
            eng.sync()
            stamps.append(time.perf_counter())
        en_cor = e
    eng.set_precision(0)
    return en_cor


def converge(eng, diis_space, tol=1e-11, max_iter=200):
    conv = [1.0, 1.0, 1.0]
    en = 0.0
    for it in range(1, max_iter + 1):
        eng.update()
        e = eng.energy()
        conv = [conv[1], conv[2], en - e]
        en = e
        eng.diis_record(it)
        eng.residuum(it)
        if it % (diis_space + 1) == 0:
            eng.diis_extrapolate()
        if all(abs(c) < tol for c in conv):
            return en, it
    return en, max_iter


def cpu_baseline(s, ints, n_iter_needed, budget_s=25.0):
    """Oracle (numpy, OpenBLAS threads) sweeps on the same system; extrapolated."""
    from oracle import solver as osolver
    o, n = s.nocc, s.norb
    t = osolver.zero_amplitudes('CCSD', o, o, n - o, n - o)
    prov = osolver.CCSDResiduals(s.h, ints.e2int, 0, o, o)
    D = osolver.block_denominators(t, ints.f_a, ints.f_b, o, o, 0.0)
    seq = osolver.sweep_order('CCSD', False)
    done = 0
    t0 = time.perf_counter()
    while True:
        osolver.sweep(t, seq, prov, D)
        osolver.energy(ints.f_a, ints.f_b, ints.v_aa, ints.v_ab, ints.v_bb, o, o, t)
        done += 1
        if time.perf_counter() - t0 > budget_s or done >= 3:
            break
    per_iter = (time.perf_counter() - t0) / done
    threads = int(os.environ.get("OPENBLAS_NUM_THREADS") or os.environ.get("OMP_NUM_THREADS")
                  or os.cpu_count() or 1)
    return dict(value=per_iter * n_iter_needed, unit="s", cores=threads, kind="port",
                sample=f"{done} oracle CCSD sweep(s)+energy of the same system "
                       f"({per_iter:.2f} s each, numpy/OpenBLAS), x {n_iter_needed} iterations")


def measured_traffic(kernel_class: str):
    """HBM bytes per launch of a GEMM class, from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs summarised under profiles/ (None if not measured)."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and \
                        parts[1] == kernel_class:
                    return float(parts[2])
    return None


def copy_probe(nbytes: int):
    """A plain device-to-device copy of a known size, three times: the
    reference reading for FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copy probe: 3 x {nbytes} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ccsd-n100", choices=sorted(
        k for k, v in synth.WORKLOADS.items() if v["method"] in ("CCSD", "CCSDT", "CCSDTQ")))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rhf", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (calibrates HBM counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding over several GPUs "
                                   "is designed but not built, so --gpus must be 1"}))
        sys.exit(2)
    diis = 5
    t0 = time.perf_counter()
    s, ints = make_workload(args.workload)
    o, n = s.nocc, s.norb
    log(f"workload built in {time.perf_counter() - t0:.1f}s: n={n} o={o} v={n - o}")
    method = synth.WORKLOADS[args.workload]["method"]
    eng = E.Engine(0, o, o, n, method, rhf=args.rhf, diis_space=diis)
    t0 = time.perf_counter()
    eng.set_ints(ints.f_a, ints.f_b, ints.v_aa, ints.v_ab, ints.v_bb)
    log(f"upload {time.perf_counter() - t0:.1f}s")
    zero = np.zeros(eng.t_size)
    # warm-up (both precisions) and the converged FP64 reference (untimed)
    eng.set_t(zero)
    run_iterations(eng, args.warmup, diis)
    eng.set_t(zero)
    run_iterations(eng, args.warmup, diis, switch_de=0.0)
    params, guess_kind = None, "mp2"
    if os.path.exists(guess.WEIGHTS):
        params, guess_kind = guess.load_weights(), "learned"
    eng.guess(guess_kind, params)
    eng.set_t(zero)
    e_conv, it_conv = converge(eng, diis)
    t_conv = eng.get_t()
    log(f"E_conv = {e_conv:.12f} after {it_conv} iterations")
    legs = {}
    for leg, switch, start_guess in (("plain", None, None), ("learned", None, guess_kind),
                                      ("mixed+learned", SWITCH_DE, guess_kind)):
        eng.set_t(zero)
        energies, stamps = [], []
        eng.sync()
        start = time.perf_counter()
        if start_guess:
            eng.guess(start_guess, params)
        run_iterations(eng, args.steps, diis, energies, stamps, switch_de=switch)
        eng.sync()
        total = time.perf_counter() - start
        hit = next((i for i, e in enumerate(energies) if abs(e - e_conv) < 1e-7), None)
        t_end = eng.get_t()
        legs[leg] = dict(total=total, value=(stamps[hit] - start) if hit is not None else None,
                         n_needed=hit + 1 if hit is not None else None,
                         e_err=abs(energies[-1] - e_conv),
                         t2_err=float(np.abs(t_end - t_conv).max()))
        log(f"{leg}: {args.steps} iterations {total * 1e3 / args.steps:.3f} ms/iter; "
            f"|E-E_conv|<1e-7 at iteration {legs[leg]['n_needed']}: {legs[leg]['value']}; "
            f"final |dE| {legs[leg]['e_err']:.2e}, max|dT| {legs[leg]['t2_err']:.2e}")
    value = legs["mixed+learned"]["value"]
    n_needed = legs["plain"]["n_needed"]
    total = legs["mixed+learned"]["total"]
    # roofline of the dominant kernel: one profiled sweep (HIP events on the stream)
    eng.profile(True)
    eng.update()
    rep = eng.profile_report()
    eng.profile(False)
    gemms = {k: v for k, v in rep.items() if k.startswith("gemm:")}
    dom = max(gemms.items(), key=lambda kv: kv[1]["ms"])
    dom_name, d = dom
    avg_s = d["ms"] / d["calls"] * 1e-3
    fl_per = d["flops"] / d["calls"]
    achieved = fl_per / avg_s
    sweep_ms = sum(v["ms"] for v in rep.values())
    gemm_ms = sum(v["ms"] for v in gemms.values())
    gemm_fl = sum(v["flops"] for v in gemms.values())
    log(f"dominant {dom_name}: {d['calls']} launches, avg {avg_s * 1e6:.1f} us, "
        f"{achieved / 1e12:.2f} TF/s; all GEMMs {gemm_fl / 1e9:.1f} GF in {gemm_ms:.2f} ms "
        f"({gemm_fl / (gemm_ms * 1e-3) / 1e12:.2f} TF/s); profiled sweep {sweep_ms:.2f} ms")
    top = sorted(rep.items(), key=lambda kv: -kv[1]["ms"])[:16]
    for k, v in top:
        log(f"  {k:36s} calls={v['calls']:4d} ms={v['ms']:.3f} avg_us={v['ms'] * 1e3 / v['calls']:.1f}")
    cpu = None
    if not args.no_cpu and n_needed is not None and method == "CCSD":
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(s, ints, n_needed)
        log(f"cpu baseline {cpu['value']:.1f} s ({cpu['sample']})")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total * 1e3 / args.steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64/f32 mixed",
        "data": "synthetic (seeded integrals, ccq onebody/twobody construction)",
        "config": {"workload": f"{args.workload}: {method}, n={n} o={o} v={n - o}, "
                               f"L={synth.WORKLOADS[args.workload]['nL']}, seed "
                               f"{synth.WORKLOADS[args.workload]['seed']}, diis 5, "
                               f"{'rhf' if args.rhf else 'uhf'} sweep",
                   "e_conv": e_conv,
                   "mode": f"mixed+learned: {guess_kind} T2 guess, FP32-MFMA tiled contractions "
                           f"until |dE|<{SWITCH_DE:g}, then FP64",
                   "legs": {k: {"time_to_1e-7_s": v["value"], "iterations_to_1e-7": v["n_needed"],
                                "ms_per_step": v["total"] * 1e3 / args.steps,
                                "final_energy_error": v["e_err"],
                                "final_max_abs_t_error": v["t2_err"]} for k, v in legs.items()},
                   "speedup_vs_plain": (legs["plain"]["value"] / value
                                        if value and legs["plain"]["value"] else None)},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64_MFMA / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_MFMA,
                     "traffic": measured_traffic(dom_name),
                     "kernel": dom_name, "avg_launch_us": avg_s * 1e6,
                     "flops_per_launch": fl_per,
                     "all_gemm_tflops": gemm_fl / (gemm_ms * 1e-3) / 1e12},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
