"""Learned T2 starting guess: data collection, GPU training and weights I/O.

The network and its kernels live in ccq_amd/csrc/guess.hip (inference,
forward / backward, Adam); this module drives them:

* ``collect``: converge CCSD with the engine on seeded synthetic systems of
  24-71 orbitals and 4-10 occupied (the regime of the 100-orbital workload,
  synth.regime_sigma) and read every T2 tuple's features and target back
  (ccq_hip_guess_data).  Integral magnitudes shrink with size in this
  regime; a network trained on 24-32 orbitals alone captures about half of
  the MP2-to-CCSD correction at 64 orbitals, the 24-71 one brings the
  100-orbital energy to 1e-6 of E_conv at the first DIIS extrapolation
  (zero start: 9e-5) and to 1e-7 at iteration 12 instead of 13;
* ``train``: normalisation from the data, then Adam on the GPU
  (ccq_hip_mlp_train);
* ``load_weights`` / ``save_weights``: the packed float32 parameter vector
  (ccq_amd/guess_weights.npz is the network this repository ships).

    python -m ccq_amd.guess --systems 500 --out ccq_amd/guess_weights.npz
"""

from __future__ import annotations

import argparse
import ctypes
import os
import sys
import time

import numpy as np

from . import engine as E
from . import synth

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "guess_weights.npz")
F, H = 5, 64


def nparams() -> int:
    return int(E.load_library().ccq_hip_mlp_nparams())


def _fptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def training_systems(count: int, seed: int = 2024, norb=(24, 72), nocc=(4, 11)):
    """(norb, nocc, nL, seed, sigma) of the seeded training systems (ranges half-open)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(*norb))
        o = int(rng.integers(*nocc))
        L = 2 * n
        out.append((n, o, L, 10000 + k, synth.regime_sigma(n, o, L)))
    return out


def make_engine(n, o, L, seed, sigma, calc_type="CCSD", **kw):
    from .inputs import build_integrals
    s = synth.make_system(n, o, L, seed, sigma=sigma)
    ints = build_integrals(s.h, s.e2int(), s.enuc, 0, o, o)
    eng = E.Engine(0, o, o, n, calc_type, **kw)
    eng.set_ints(ints.f_a, ints.f_b, ints.v_aa, ints.v_ab, ints.v_bb)
    return eng


def collect(systems, tol=1e-10, log=None):
    """Features / targets of every T2 tuple of the converged systems."""
    from .solve import jacobi_iter
    X, M, Y = [], [], []
    for k, spec in enumerate(systems):
        eng = make_engine(*spec)
        r = jacobi_iter(eng, tol, 100, 5, out=None)
        if r.converged:
            x, m, y = eng.guess_data()
            X.append(x)
            M.append(m)
            Y.append(y)
        eng.close()
        if log and (k + 1) % 50 == 0:
            log(f"collected {k + 1} systems")
    return np.concatenate(X), np.concatenate(M), np.concatenate(Y)


def init_params(X, seed=0):
    rng = np.random.default_rng(seed)
    mu = X.mean(0).astype(np.float64)
    sd = X.std(0).astype(np.float64) + 1e-6
    parts = [mu, sd, rng.normal(0, 1 / np.sqrt(F), F * H), np.zeros(H),
             rng.normal(0, 1 / np.sqrt(H), H * H), np.zeros(H), rng.normal(0, 1 / np.sqrt(H), H),
             np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def train(X, M, Y, params, steps=3000, batch=4096, lr=3e-3, seed=1):
    lib = E.load_library()
    X = np.ascontiguousarray(X, dtype=np.float32)
    M = np.ascontiguousarray(M, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    p = np.ascontiguousarray(params, dtype=np.float32).copy()
    loss = ctypes.c_double()
    rc = lib.ccq_hip_mlp_train(_fptr(X), _fptr(M), _fptr(Y), len(M), _fptr(p), p.size, steps, batch,
                               lr, seed, ctypes.byref(loss))
    if rc != 0:
        raise E.EngineError("ccq_hip_mlp_train: " + lib.ccq_hip_mlp_last_error().decode())
    return p, loss.value


def grad(X, M, Y, params, noalloc=False):
    """Loss and gradient over all samples; noalloc selects the entry point
    that uses the library's static per-GPU scratch instead of allocating."""
    lib = E.load_library()
    X = np.ascontiguousarray(X, dtype=np.float32)
    M = np.ascontiguousarray(M, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    p = np.ascontiguousarray(params, dtype=np.float32)
    g = np.zeros(p.size, dtype=np.float32)
    loss = ctypes.c_double()
    name = "ccq_hip_mlp_grad_noalloc" if noalloc else "ccq_hip_mlp_grad"
    rc = getattr(lib, name)(_fptr(X), _fptr(M), _fptr(Y), len(M), _fptr(p), p.size, _fptr(g),
                            ctypes.byref(loss))
    if rc != 0:
        raise E.EngineError(f"{name}: " + lib.ccq_hip_mlp_last_error().decode())
    return loss.value, g


def save_weights(path, params, meta):
    np.savez(path, params=np.asarray(params, dtype=np.float32), **meta)


def load_weights(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--systems", type=int, default=500)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    log = lambda m: print(f"[guess] {m}", flush=True)
    t0 = time.perf_counter()
    X, M, Y = collect(training_systems(args.systems), log=log)
    log(f"{len(M)} samples from {args.systems} systems in {time.perf_counter() - t0:.1f}s")
    p0 = init_params(X)
    t0 = time.perf_counter()
    p, loss = train(X, M, Y, p0, args.steps, args.batch, args.lr)
    log(f"trained {args.steps} steps in {time.perf_counter() - t0:.1f}s, batch loss {loss:.4e}, "
        f"target variance {float(np.mean(Y.astype(np.float64) ** 2)):.4e}")
    save_weights(args.out, p, dict(systems=args.systems, steps=args.steps, batch=args.batch,
                                   lr=args.lr, final_loss=loss))
    log(f"wrote {args.out}")


if __name__ == "__main__":
    sys.exit(main())
