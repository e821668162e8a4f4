"""Pins of the spin-orbital CCSDT oracle (oracle/ccsdt_so.py).

Every residual element of R1, R2 and R3 must equal the brute-force
determinant-space projection <mu|exp(-T) H exp(T)|0> (oracle/bruteforce.py)
at random antisymmetric T1, T2, T3, closed and open shell, with and without
T3 (the T3-free case is CCSD's residual plus the CCSD moment M3).  A wrong
coefficient, sign or index pairing in any term changes these elements by
O(|T|^2) and fails the 1e-12 bar.
"""

import numpy as np
import pytest

from oracle import bruteforce as bf
from oracle import ccsdt_so as so
from oracle import solver as osolver
from ccq_amd import synth


def _random_so(space, rng, rank, scale=0.1):
    No, Nv = space.nocc_so, space.nvir_so
    x = scale * rng.standard_normal((No,) * rank + (Nv,) * rank)
    x = so._anti(x, [list(range(rank)), list(range(rank, 2 * rank))])
    # keep only spin-allowed elements
    return bf.blocks_to_so(space, rank, bf.so_to_blocks(space, rank, x))


@pytest.mark.parametrize("n,oa,ob,seed", [(6, 3, 3, 11), (7, 3, 2, 12), (6, 2, 1, 13)])
def test_residuals_match_bruteforce(n, oa, ob, seed):
    s = synth.make_system(n, oa, 20, seed)
    space = bf.DetSpace(n, oa, ob)
    H = bf.hamiltonian(space, s.h, s.eri)
    rng = np.random.default_rng(seed)
    t1, t2, t3 = (_random_so(space, rng, k) for k in (1, 2, 3))
    res = so.Residuals(s.h, s.eri, oa, ob)
    for amps in ({1: t1, 2: t2, 3: t3}, {1: t1, 2: t2}):
        want, _ = bf.residuals(space, H, amps, [1, 2, 3])
        got = res.residuals(t1, t2, amps.get(3))
        for k in (1, 2, 3):
            assert np.abs(want[k]).max() > 1e-2
            np.testing.assert_allclose(got[k], want[k], atol=1e-12, rtol=0, err_msg=f"R{k}")


def test_block_residuals_drive_ccq_sweep():
    """The block provider reproduces the brute-force provider inside ccq's
    in-place sweep (src/update_cc.F90:102-318), open shell, uhf order."""
    n, oa, ob = 6, 3, 2
    s = synth.make_system(n, oa, 20, 21)
    from oracle import integrals as I
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, oa, ob)
    rng = np.random.default_rng(4)
    t = osolver.zero_amplitudes('CCSDT', oa, ob, n - oa, n - ob)
    space = bf.DetSpace(n, oa, ob)
    for name in t:
        rank, nb = osolver.BLOCKS[name]
        t[name] = bf.so_to_blocks(space, rank, _random_so(space, rng, rank, 0.05))[nb]
    t_ref = {k: v.copy() for k, v in t.items()}
    D = osolver.block_denominators(t, fa, fb, oa, ob, 0.01)
    seq = osolver.sweep_order('CCSDT', False)
    osolver.sweep(t, seq, so.SweepResiduals(s.h, s.eri, oa, ob), D)
    osolver.sweep(t_ref, seq, osolver.BruteForceResiduals(s.h, s.eri, n, oa, ob, 3), D)
    for k in t:
        np.testing.assert_allclose(t[k], t_ref[k], atol=1e-12, rtol=0, err_msg=k)
