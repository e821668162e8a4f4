"""Write tests/golden/ccsdtq_sweep_n8.npz: one CCSDTQ Gauss-Seidel sweep of
the brute-force determinant oracle (ccq's update order, src/update_cc.F90:
12-323) on an 8-orbital system with 4 alpha and 3 beta electrons, where every
T4 spin block that can be non-zero is (t4aaaa needs 4 occupied alpha).

The oracle sweep takes a few minutes, so its result is stored: the canonical
amplitudes (strictly increasing inside each antisymmetric group) of every
block after the sweep.  Inputs are regenerated from seeds by
gpu_helpers.ccsdtq_sweep_case.

    python tests/golden/make_ccsdtq_sweep.py
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from oracle import solver as osolver  # noqa: E402
import gpu_helpers as G  # noqa: E402


def main():
    case = G.ccsdtq_sweep_case()
    t = {k: v.copy() for k, v in case["t"].items()}
    prov = osolver.BruteForceResiduals(case["h"], case["eri"], case["n"], case["oa"], case["ob"], 4)
    D = osolver.block_denominators(t, case["fa"], case["fb"], case["oa"], case["ob"], case["shift"])
    t0 = time.time()
    osolver.sweep(t, osolver.sweep_order("CCSDTQ", False), prov, D)
    print(f"oracle sweep {time.time() - t0:.1f}s")
    out = {}
    for k, x in t.items():
        out[k] = x[G.canonical_mask(k, x.shape)]
    np.savez_compressed(os.path.join(HERE, "ccsdtq_sweep_n8.npz"), **out)
    print({k: v.size for k, v in out.items()})


if __name__ == "__main__":
    main()
