"""Benchmark: time-to-|dE|<1e-7 Eh of the CC iteration on one MI355X,
mixed precision plus the learned guess against plain FP64.

Workload (N=1): CCSDT, 40 spatial orbitals (5 occupied, 35 virtual),
synthetic integrals (pq|rs) = sum_L B B with L = 80 from seed 2 (B entries
N(0, 0.05^2)), ccq's iteration (diis_space 5, shift 0) with rhf=.true. --
the closed-shell sweep ccq's own CCSDT / CCSDTQ inputs use
(tests/h4_sto3g/CCSDT.inp, tests/f2_vdz_2.0/CCSDT.inp).  The same legs with
the uhf sweep (rhf=.false.) are reported in config.uhf (--no-uhf-legs skips
them; --uhf swaps the two).  A second workload -- CCSD, 100 orbitals / 10
occupied, L = 200, seed 1, B entries N(0, 0.01137^2) -- is measured the same
way and reported in config.secondary (--secondary none skips it).

Protocol (per workload)
  1. build the system and upload it (untimed);
  2. W warm-up iterations from T = 0 in each precision, then a tight untimed
     FP64 solve (tol 1e-11) for the converged energy E_conv and amplitudes;
  3. plain leg: T = 0, FP64 iterations (update, energy, DIIS ring, residuum,
     DIIS every 6th), stream synchronised after every iteration: exactly K
     of them (ms_per_step), continued until the first iteration within
     1e-7 of E_conv;
  4. learned leg: the learned starting guess (computed inside the timed
     region; MP2 if no weights file), then FP64 iterations as above;
  5. mixed+learned leg: the learned guess, then FP32 sweeps (FP32-MFMA
     contractions on FP32 copies of every tensor of the sweep; the
     amplitudes themselves stay FP64) until |dE| < 3e-5 (above the FP32
     noise of dE on these systems), FP64 after
     that; the iteration of the switch is recorded and, as in solve.Mixed,
     only FP64 sweeps count as converged;
  6. value = the mixed+learned leg's wall time until the first FP64
     iteration with |E - E_conv| < 1e-7.
  Roofline: the contraction class with the largest share of one profiled
  FP64 sweep (HIP events on the engine's stream): 2*M*N*K flops per launch
  over its average launch time.  traffic: HBM bytes per launch of that class
  from the rocprofv3 FETCH_SIZE / WRITE_SIZE summaries under profiles/
  (copy-calibrated), when one is recorded for it.
  cpu_baseline: the CPU oracle (numpy / OpenBLAS on the host cores) on the
  same system -- CCSDT: one evaluation of every residual block by the
  spin-orbital oracle plus the Jacobi step; CCSD: the spin-integrated oracle
  sweep -- timed and multiplied by the plain leg's iteration count.

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
# tile shapes and split-K from the measuring autotuner (set CCQ_HIP_TUNE=0 for the model's plans)
os.environ.setdefault("CCQ_HIP_TUNE", "1")

from ccq_amd import engine as E  # noqa: E402
from ccq_amd import guess  # noqa: E402
from ccq_amd import synth  # noqa: E402
from ccq_amd.inputs import build_integrals  # noqa: E402

METRIC = "time-to-|dE|<1e-7 Eh, mixed+learned vs plain FP64 (1 GPU); % FP64 MFMA peak"
PEAK_FP64_MFMA = 78.6e12
SWITCH_DE = 3e-5     # mixed leg: FP32 contractions until |dE| falls below this
MIXED_MODE = 2       # precision mode of the FP32 sweeps: FP32 MFMA on FP32 storage
HIT = 1e-7
MAX_ITER = 200       # a leg that has not reached HIT after K iterations continues up to this


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_desc(name, rhf):
    w = synth.WORKLOADS[name]
    sigma = w.get("sigma", 0.05)
    n, o = w["norb"], w["nocc"]
    return (f"{name}: {w['method']}, n={n} o={o} v={n - o}, L={w['nL']}, seed {w['seed']}, "
            f"B ~ N(0, {sigma:.5f}^2), diis 5, {'rhf' if rhf else 'uhf'} sweep")


class Leg:
    """One timed run of ccq's loop from a fresh start."""

    def __init__(self, eng, steps, diis, e_conv, t_conv):
        self.eng, self.steps, self.diis = eng, steps, diis
        self.e_conv, self.t_conv = e_conv, t_conv

    def run(self, guess_kind=None, params=None, switch_de=None):
        eng = self.eng
        eng.set_t(np.zeros(eng.t_size))
        if eng.t4_size:
            eng.set_t4(np.zeros(eng.t4_size))
        energies, stamps, fp64 = [], [], []
        low = switch_de is not None
        switch_it = None
        eng.set_precision(MIXED_MODE if low else 0)
        en_cor = 0.0
        eng.sync()
        start = time.perf_counter()
        if guess_kind:
            eng.guess(guess_kind, params)
        it, reached = 0, False
        # K iterations, continued (up to MAX_ITER) until the first FP64 sweep within HIT
        while it < self.steps or (not reached and it < MAX_ITER):
            it += 1
            fp64.append(not low)
            eng.update()
            e = eng.energy()
            if low and abs(en_cor - e) < switch_de:
                low = False
                switch_it = it
                eng.set_precision(0)
            eng.diis_record(it)
            eng.residuum(it)
            if it % (self.diis + 1) == 0:
                eng.diis_extrapolate()
            eng.sync()
            stamps.append(time.perf_counter())
            energies.append(e)
            en_cor = e
            reached = reached or (fp64[-1] and abs(e - self.e_conv) < HIT)
        eng.set_precision(0)
        total = stamps[self.steps - 1] - start      # exactly K iterations
        hit = next((i for i, e in enumerate(energies)
                    if fp64[i] and abs(e - self.e_conv) < HIT), None)
        t_end = eng.get_t()
        return dict(total=total, value=(stamps[hit] - start) if hit is not None else None,
                    n_needed=hit + 1 if hit is not None else None, switch=switch_it,
                    e_err=abs(energies[-1] - self.e_conv),
                    t_err=float(np.abs(t_end - self.t_conv).max()),
                    fp32_sweeps=int(sum(1 for f in fp64 if not f)), iterations=it)


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


def cpu_threads():
    return int(os.environ.get("OPENBLAS_NUM_THREADS") or os.environ.get("OMP_NUM_THREADS")
               or os.cpu_count() or 1)


def cpu_baseline(s, ints, method, t_vec, n_iter):
    """The CPU oracle on the same system, one sweep's worth of residual work,
    times the plain leg's iteration count."""
    from oracle import solver as osolver
    o, n = s.nocc, s.norb
    v = n - o
    t = E.split_tvec(t_vec, o, o, v, v)
    D = osolver.block_denominators(t, ints.f_a, ints.f_b, o, o, 0.0)
    if method == "CCSDT":
        from oracle import ccsdt_so
        res = ccsdt_so.Residuals(s.h, s.eri, o, o)
        t0 = time.perf_counter()
        R = res.block_residuals(t)
        for k in R:
            t[k] = t[k] - R[k] / D[k]
        osolver.energy(ints.f_a, ints.f_b, ints.v_aa, ints.v_ab, ints.v_bb, o, o, t)
        per = time.perf_counter() - t0
        sample = (f"1 residual evaluation of every T1-T3 block by the spin-orbital CCSDT oracle "
                  f"(oracle/ccsdt_so.py, numpy/OpenBLAS) + Jacobi step + energy on the same "
                  f"system ({per:.1f} s), x {n_iter} iterations")
    else:
        prov = osolver.CCSDResiduals(s.h, ints.e2int, 0, o, o)
        seq = osolver.sweep_order('CCSD', False)
        t0 = time.perf_counter()
        osolver.sweep(t, seq, prov, D)
        osolver.energy(ints.f_a, ints.f_b, ints.v_aa, ints.v_ab, ints.v_bb, o, o, t)
        per = time.perf_counter() - t0
        sample = (f"1 spin-integrated CCSD oracle sweep + energy of the same system "
                  f"({per:.1f} s, numpy/OpenBLAS), x {n_iter} iterations")
    return dict(value=per * n_iter, unit="s", cores=cpu_threads(), kind="port", sample=sample)


def measured_traffic(kernel_class: str):
    """HBM bytes per launch of a kernel class from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE summaries under profiles/ (None if not measured).  A summary
    names the class with its plan (gemm:MxNxK/plan) or by its shape alone
    (gemm:MxNxK, for plans chosen by the autotuner); the exact name wins."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    shape = kernel_class.split("/")[0]
    found = {}
    for name in sorted(os.listdir(pdir)):
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch":
                    found[parts[1]] = float(parts[2])
    return found.get(kernel_class, found.get(shape))


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


def mfma_ceiling(iters=2000):
    """The attainable MFMA rate of this part (pure MFMA loops, 8 independent
    accumulators per wave) for FP64, FP32 and BF16, at 1, 2 and 4 waves per
    SIMD (256 CUs x 4 SIMDs); one JSON line per point."""
    import ctypes
    lib = E.load_library()
    peaks = {0: ("f64 16x16x4", 78.6e12), 1: ("f32 16x16x4", 157.3e12), 2: ("bf16 16x16x32", 2.5e15)}
    for kind, (name, peak) in peaks.items():
        for blocks in (256, 512, 1024):
            sec, fl = ctypes.c_double(), ctypes.c_double()
            it = iters if kind < 2 else iters // 4
            rc = lib.ccq_hip_mfma_ceiling(kind, blocks, it, ctypes.byref(sec), ctypes.byref(fl))
            if rc != 0:
                raise E.EngineError(f"ccq_hip_mfma_ceiling: HIP error {rc}")
            rate = fl.value / sec.value
            print(json.dumps({"mfma": name, "workgroups": blocks, "waves_per_simd": blocks // 256,
                              "seconds": sec.value, "tflops": rate / 1e12, "peak_tflops": peak / 1e12,
                              "frac": rate / peak}), flush=True)


def sweeps_only(name, count, prec, rhf):
    """Only `count` sweeps (graphed, as in the timed legs) from the state after
    six ccq iterations: a short, representative program for profilers."""
    method = synth.WORKLOADS[name]["method"]
    s = synth.workload_system(name)
    ints = build_integrals(s.h, s.e2int(), s.enuc, 0, s.nocc, s.nocc)
    eng = E.Engine(0, s.nocc, s.nocc, s.norb, method, rhf=rhf, diis_space=5)
    eng.set_ints(ints.f_a, ints.f_b, ints.v_aa, ints.v_ab, ints.v_bb)
    for it in range(1, 7):
        eng.update()
        eng.energy()
        eng.diis_record(it)
        eng.residuum(it)
    eng.set_precision(prec)
    eng.update()   # graph capture (and, in mode 2, the float shadows) untimed
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(count):
        eng.update()
    eng.sync()
    log(f"{name}: {count} sweeps (precision {prec}) in {(time.perf_counter() - t0) * 1e3:.2f} ms")
    eng.close()


def measure(name, steps, warmup, rhf, with_cpu, profile=True):
    diis = 5
    method = synth.WORKLOADS[name]["method"]
    t0 = time.perf_counter()
    s = synth.workload_system(name)
    ints = build_integrals(s.h, s.e2int(), s.enuc, 0, s.nocc, s.nocc)
    o, n = s.nocc, s.norb
    log(f"{name}: system built in {time.perf_counter() - t0:.1f}s (n={n} o={o} v={n - o})")
    eng = E.Engine(0, o, o, n, method, rhf=rhf, diis_space=diis)
    eng.set_ints(ints.f_a, ints.f_b, ints.v_aa, ints.v_ab, ints.v_bb)
    params, guess_kind = None, "mp2"
    if os.path.exists(guess.WEIGHTS):
        params, guess_kind = guess.load_weights(), "learned"
    # warm-up (both precisions, the guess kernels) and the converged FP64 solution
    warm = Leg(eng, warmup, diis, 0.0, 0.0)
    warm.t_conv = np.zeros(eng.t_size)
    warm.run()
    warm.run(guess_kind, params, switch_de=0.0)
    eng.set_t(np.zeros(eng.t_size))
    if eng.t4_size:
        eng.set_t4(np.zeros(eng.t4_size))
    e_conv, it_conv = converge(eng, diis)
    t_conv = eng.get_t()
    log(f"{name}: E_conv = {e_conv:.12f} after {it_conv} iterations")
    leg = Leg(eng, steps, diis, e_conv, t_conv)
    legs = {}
    for label, g, sw in (("plain", None, None), ("learned", guess_kind, None),
                         ("mixed+learned", guess_kind, SWITCH_DE)):
        r = leg.run(g, params, sw)
        legs[label] = r
        log(f"{name} {label}: {steps} iterations {r['total'] * 1e3 / steps:.3f} ms/iter; "
            f"|E-E_conv|<1e-7 at iteration {r['n_needed']} ({r['value']} s); switch at "
            f"{r['switch']}; final |dE| {r['e_err']:.2e}, max|dT| {r['t_err']:.2e}")
    out = dict(workload=workload_desc(name, rhf), method=method, e_conv=e_conv, legs=legs,
               guess=guess_kind)
    if profile:
        eng.set_t(t_conv)
        eng.profile(True)
        eng.update()
        rep = eng.profile_report()
        eng.profile(False)
        gemms = {k: v for k, v in rep.items() if k.startswith("gemm:")}
        dom_name, d = max(gemms.items(), key=lambda kv: kv[1]["ms"])
        avg_s = d["ms"] / d["calls"] * 1e-3
        fl_per = d["flops"] / d["calls"]
        sweep_ms = sum(v["ms"] for v in rep.values())
        gemm_ms = sum(v["ms"] for v in gemms.values())
        gemm_fl = sum(v["flops"] for v in gemms.values())
        log(f"{name}: dominant {dom_name}: {d['calls']} launches, avg {avg_s * 1e6:.1f} us, "
            f"{fl_per / avg_s / 1e12:.2f} TF/s; all GEMMs {gemm_fl / 1e9:.1f} GF in "
            f"{gemm_ms:.2f} ms ({gemm_fl / (gemm_ms * 1e-3) / 1e12:.2f} TF/s); profiled sweep "
            f"{sweep_ms:.2f} ms")
        for k, v in sorted(rep.items(), key=lambda kv: -kv[1]["ms"])[:24]:
            log(f"  {k:40s} calls={v['calls']:4d} ms={v['ms']:.3f} "
                f"avg_us={v['ms'] * 1e3 / v['calls']:.1f}")
        out["roofline"] = dict(kernel=dom_name, avg_s=avg_s, flops=fl_per, sweep_ms=sweep_ms,
                               gemm_tflops=gemm_fl / (gemm_ms * 1e-3) / 1e12, gemm_gf=gemm_fl / 1e9)
    eng.close()
    if with_cpu and legs["plain"]["n_needed"]:
        log(f"{name}: cpu baseline (oracle) ...")
        out["cpu"] = cpu_baseline(s, ints, method, t_conv, legs["plain"]["n_needed"])
        log(f"{name}: cpu baseline {out['cpu']['value']:.1f} s ({out['cpu']['sample']})")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ccsdt-n40", choices=sorted(
        k for k, v in synth.WORKLOADS.items() if v["method"] in ("CCSD", "CCSDT", "CCSDTQ")))
    ap.add_argument("--secondary", default="ccsd-n100",
                    help="second workload reported in config.secondary, or 'none'")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--uhf", action="store_true",
                    help="uhf sweeps (rhf=.false.) for the timed legs; default rhf=.true., as ccq's "
                         "closed-shell CCSDT / CCSDTQ inputs run")
    ap.add_argument("--no-uhf-legs", action="store_true",
                    help="skip the uhf legs of the main workload reported in config.uhf")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (calibrates HBM counters)")
    ap.add_argument("--sweeps-only", type=int, default=0, metavar="N",
                    help="only run N graphed sweeps of --workload (for profilers)")
    ap.add_argument("--precision", type=int, default=0, help="with --sweeps-only: 0 FP64, 1 FP32 MFMA, 2 FP32 MFMA and storage")
    ap.add_argument("--mfma-ceiling", action="store_true",
                    help="only measure the attainable MFMA rates (pure MFMA loops)")
    args = ap.parse_args()
    if args.mfma_ceiling:
        mfma_ceiling()
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.sweeps_only:
        sweeps_only(args.workload, args.sweeps_only, args.precision, not args.uhf)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding over several GPUs "
                                   "is designed but not built, so --gpus must be 1"}))
        sys.exit(2)
    rhf = not args.uhf
    main_r = measure(args.workload, args.steps, args.warmup, rhf, not args.no_cpu)
    other = None
    if not args.no_uhf_legs:
        other = measure(args.workload, args.steps, args.warmup, not rhf, False, profile=False)
    sec = None
    if args.secondary and args.secondary != "none" and args.secondary != args.workload:
        sec = measure(args.secondary, args.steps, args.warmup, rhf, False, profile=True)
    legs = main_r["legs"]
    value = legs["mixed+learned"]["value"]
    rf = main_r["roofline"]
    achieved = rf["flops"] / rf["avg_s"]

    def summary(r):
        lg = r["legs"]
        out = {"workload": r["workload"], "e_conv": r["e_conv"],
               "legs": {k: {"time_to_1e-7_s": v["value"], "iterations_to_1e-7": v["n_needed"],
                            "ms_per_step": v["total"] * 1e3 / args.steps,
                            "switch_to_fp64_at": v["switch"], "fp32_sweeps": v["fp32_sweeps"],
                            "final_energy_error": v["e_err"], "final_max_abs_t_error": v["t_err"]}
                        for k, v in lg.items()},
               "speedup_vs_plain": (lg["plain"]["value"] / lg["mixed+learned"]["value"]
                                    if lg["mixed+learned"]["value"] and lg["plain"]["value"]
                                    else None)}
        if "roofline" in r:
            out["profiled_sweep_ms"] = r["roofline"]["sweep_ms"]
            out["all_gemm_tflops"] = r["roofline"]["gemm_tflops"]
        return out

    cfg = summary(main_r)
    cfg["mode"] = (f"mixed+learned: {main_r['guess']} starting guess, FP32 sweeps (precision mode "
                   f"{MIXED_MODE}: FP32-MFMA contractions on FP32 storage) until "
                   f"|dE|<{SWITCH_DE:g}, then FP64; only FP64 sweeps count as converged")
    if other:
        cfg["uhf" if rhf else "rhf"] = summary(other)
    if sec:
        cfg["secondary"] = summary(sec)
    cpu = main_r.get("cpu")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": legs["mixed+learned"]["total"] * 1e3 / args.steps,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (plain), f32/f64 (mixed)",
        "data": "synthetic (seeded integrals, ccq onebody/twobody construction)",
        "config": cfg,
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12,
                     "peak": PEAK_FP64_MFMA / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_MFMA, "traffic": measured_traffic(rf["kernel"]),
                     "kernel": rf["kernel"], "avg_launch_us": rf["avg_s"] * 1e6,
                     "flops_per_launch": rf["flops"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
