"""Write tests/golden/cct3_n7_<oa><ob>.npz: the CC(t;3) tail of the oracle
(oracle/cct3.py) on converged active-space CCSDt amplitudes of a 7-orbital
system (act_occ = act_unocc = 1), closed shell (3, 3) and open shell (3, 2).

The brute-force CCSDt solve takes about half a minute per case, so its result
is stored with the tail's outputs:

  tvec      ccq's packed CCSDt t_vec at convergence (tol 1e-11)
  ecor      its correlation energy
  l_vec     the converged left-CCSD vector (T1/T2 layout of t_vec)
  left_e    pseudo-energy per left iteration (tol 1e-11, DIIS space 5)
  e23       crcc23's (A, B, C, D)
Inputs are regenerated from seeds by ccq_amd.synth.make_system(7, 3, 20, 11).

    python tests/golden/make_cct3.py
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import active as A  # noqa: E402
from oracle import bruteforce as bf  # noqa: E402
from oracle import cct3 as C  # noqa: E402
from oracle import integrals as I  # noqa: E402
from oracle import solver as S  # noqa: E402
from ccq_amd import synth  # noqa: E402

N, AO, AU, SEED = 7, 1, 1, 11
TOL, DIIS = 1e-11, 5


def case(oa, ob):
    s = synth.make_system(N, oa, 20, SEED)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    act = A.ActiveSpace(0, oa, ob, N, AO, AU)
    va, vb = N - oa, N - ob
    t = S.zero_amplitudes('CCSDT', oa, ob, va, vb)
    prov = S.BruteForceResiduals(s.h, s.eri, N, oa, ob, 3)
    t0 = time.time()
    r = A.jacobi_iter(t, S.Options(calc_type='CCSDT', tol=TOL, max_iter=80, diis_space=DIIS),
                      prov, fa, fb, vaa, vab, vbb, act)
    assert r.converged
    space = bf.DetSpace(N, oa, ob)
    out = C.cct3(space, s.h, s.eri, t, act, tol=TOL, diis_space=DIIS)
    assert out['left_converged']
    print(f"case {oa},{ob}: ecor {r.en_cor:.12f} e23 {out['e23']} ({time.time() - t0:.0f}s)")
    b1 = bf.so_to_blocks(space, 1, out['l1'])
    b2 = bf.so_to_blocks(space, 2, out['l2'])
    l_vec = np.concatenate([b1[0].ravel(), b1[1].ravel(), b2[0].ravel(), b2[1].ravel(),
                            b2[2].ravel()])
    np.savez_compressed(os.path.join(HERE, f"cct3_n7_{oa}{ob}.npz"),
                        tvec=A.tvec(t, act, oa, ob, va, vb), ecor=r.en_cor, l_vec=l_vec,
                        left_e=np.array([h[1] for h in out['left_history']]),
                        e23=np.array(out['e23']))


if __name__ == "__main__":
    case(3, 3)
    case(3, 2)
