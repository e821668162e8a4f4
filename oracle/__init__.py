"""CPU oracle for the ccq coupled-cluster residual-and-update iteration.

TEST INFRASTRUCTURE ONLY.  This package is the checker: it restates, in plain
numpy, the algorithm of piecuch-group/ccq for the CC amplitude iteration
(input layout, Fock build, HF energy, the in-place block update sweep, the
Jacobi step, the correlation energy, DIIS and the stop rules), and holds a
brute-force determinant-space coupled-cluster solver that pins the residual
equations independently of any factorisation.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product (``ccq_amd``) never does.

Modules
  integrals   -- load_ints / gen_fock_operator / calc_hf_energy restatement
  bruteforce  -- determinant-space H, exp(T), projective residuals, FCI
  ccsd        -- spin-integrated CCSD residual blocks (T1-dressed form)
  solver      -- ccq's jacobi_iter: sweep order, Jacobi step, energy, DIIS,
                 convergence and print formats

Pinning status (what fixes each function besides the code itself)
  integrals.load_*        golden H4 files (tools/ccq_py/*.inp), E_HF KAT
  integrals.fock/E_HF     closed form + H4 KAT (tests/test_small.py:42)
  bruteforce.*            FCI = exact diagonalisation; CCSDTQ == FCI for 4 e-;
                          CCSD == FCI for 2 e-; H4 CCSDT/CCSDTQ KATs
  ccsd.*                  residual == brute-force <mu|exp(-T) H exp(T)|0> at
                          random T; iteration-1 T2 == MP2 amplitudes
  solver.diis             numpy least-squares restatement on random vectors
  solver.energy           closed form vs brute force <0|Hbar|0>
"""
