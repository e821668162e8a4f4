"""Pins for the CC(t;3) tail oracle (oracle/cct3.py).

* H-bar and the left equations: at converged CCSD amplitudes the left
  eigenvector of the brute-force similarity-transformed Hamiltonian (in the
  singles + doubles space) zeroes the oracle's connected left-CCSD residuals,
  and the oracle's Jacobi solver (update_l2 with H-bar denominators,
  src/update_lcc.f90:10-104, :208-280) converges to it;
* l3 = <0|(L1 + L2) Hbar|ijk abc> equals the brute-force projection for any
  T1, T2, L1, L2;
* cal_d3 (src/mm_correct.f90:345-481): closed-shell spin symmetry of the eight
  arrays, with d3d1 equal to d3a1 once its loop runs over every virtual, and
  the loop-start quirk (orbital occ_a + 2, :459) dropping exactly the first
  beta virtual(s);
* denominators: with T = 0, -D_C is the brute-force diagonal
  <ijk abc|H|ijk abc> - <0|H|0> and D_A = D_B, D_D = D_C;
* crcc23's sums: closed-shell spin symmetry case by case (aaa = bbb,
  aab = abb) and no correction when every triple is in the P space.

* D_D with three-body terms: with T1 = 0 it equals minus the brute-force
  diagonal <ijk abc|Hbar|ijk abc> - <0|Hbar|0> for every triple (the loop
  quirk switched off), D_C does not;
* the left equations at arbitrary (CCSDt) T1, T2 and L: the exact
  <0|(1+L)(Hbar-E)|mu> minus the oracle residual is zero on doubles and the
  disconnected sum_jb l_ij^ab R_j^b on singles.
"""

import itertools

import numpy as np
import pytest

from oracle import bruteforce as bf
from oracle import cct3 as C
from ccq_amd import synth


def _ccsd_state(n, oa, ob, seed=5):
    s = synth.make_system(n, oa, 20, seed)
    space = bf.DetSpace(n, oa, ob)
    no = space.nocc_so
    hso, vso = C.so_integrals(space, s.h, s.eri)
    f = C.fock(hso, vso, no)
    Hm = bf.hamiltonian(space, s.h, s.eri)
    amps, ec = bf.solve_cc(space, Hm, 2, np.diag(f)[:no], np.diag(f)[no:])
    return s, space, hso, vso, f, Hm, amps, ec


def _similarity(space, Hm, amps):
    T = bf.t_operator(space, amps).toarray()
    dim = len(space.dets)

    def expm(M):
        out, term = np.eye(dim), np.eye(dim)
        for k in range(1, 40):
            term = term @ M / k
            if not term.any():
                break
            out = out + term
        return out
    return expm(-T) @ Hm @ expm(T)


def _sd_basis(space):
    basis = []
    for rank in (1, 2):
        for occ, vir in bf.excitations(space, rank):
            sgn, d = bf._apply(bf._exc_ops(space, occ, vir), space.ref)
            basis.append((rank, occ, vir, sgn, space.index[d]))
    return basis


def _brute_lambda(space, Hb, no, nv):
    """Left eigenvector of Hbar in the S+D space: lam (Hbar_SD - E) = -Hbar_0,SD."""
    r0 = space.index[space.ref]
    basis = _sd_basis(space)
    m = len(basis)
    A = np.zeros((m, m))
    b = np.zeros(m)
    for mu, (_, _, _, sm, dm) in enumerate(basis):
        b[mu] = Hb[r0, dm] * sm
        for nu, (_, _, _, sn, dn) in enumerate(basis):
            A[mu, nu] = sn * sm * Hb[dn, dm]
        A[mu, mu] -= Hb[r0, r0]
    lam = np.linalg.solve(A, -b)
    l1 = np.zeros((no, nv))
    l2u = {}
    for x, (rk, oc, vi, _, _) in zip(lam, basis):
        if rk == 1:
            l1[oc[0], vi[0]] = x
        else:
            l2u[(oc, vi)] = x
    return l1, bf.antisym_fill(2, l2u, no, nv), lam, basis


@pytest.mark.parametrize("n,oa,ob", [(6, 3, 2), (6, 2, 2)])
def test_left_equations_vanish_at_brute_force_left_eigenvector(n, oa, ob):
    s, space, hso, vso, f, Hm, amps, ec = _ccsd_state(n, oa, ob)
    no, nv = space.nocc_so, space.nvir_so
    t1, t2 = amps[1], amps[2]
    ht, gt = C.dress(hso, vso, t1)
    H = C.Hbar(ht, gt, t2, no)
    Hb = _similarity(space, Hm, amps)
    l1b, l2b, _, _ = _brute_lambda(space, Hb, no, nv)
    assert np.abs(l2b).max() > 1e-3
    r1, r2 = C.lambda_residual(H, t2, l1b, l2b)
    assert np.abs(r1).max() < 1e-10 and np.abs(r2).max() < 1e-10
    # a dropped term is seen: without the Gae / Gmi three-body pieces the residual is not zero
    r1x, _ = C.lambda_residual(H, 0 * t2, l1b, l2b)
    assert np.abs(r1x).max() > 1e-6
    # the Jacobi solver with H-bar denominators reaches the same L
    l1, l2, ok, hist = C.solve_left(space, H, t2, ec, tol=1e-12, max_iter=200)
    assert ok
    assert np.abs(l1 - l1b).max() < 1e-8 and np.abs(l2 - l2b).max() < 1e-8
    # the pseudo-energy |LH| + E_cor tends to E_cor
    assert abs(hist[-1][1] - ec) < 1e-8


def test_l3_is_brute_force_projection():
    n, oa, ob = 6, 3, 2
    s = synth.make_system(n, oa, 20, 7)
    space = bf.DetSpace(n, oa, ob)
    no, nv = space.nocc_so, space.nvir_so
    rng = np.random.default_rng(3)

    def rnd(rank, scale):
        u = {occ + vir: scale * rng.standard_normal() for occ, vir in bf.excitations(space, rank)}
        return bf.antisym_fill(rank, {(k[:rank], k[rank:]): v for k, v in u.items()}, no, nv)
    t1, t2 = rnd(1, 0.05), rnd(2, 0.05)
    l1, l2 = rnd(1, 0.1), rnd(2, 0.1)
    hso, vso = C.so_integrals(space, s.h, s.eri)
    ht, gt = C.dress(hso, vso, t1)
    H = C.Hbar(ht, gt, t2, no)
    Hb = _similarity(space, bf.hamiltonian(space, s.h, s.eri), {1: t1, 2: t2})
    Lv = np.zeros(len(space.dets))
    for rank, l in ((1, l1), (2, l2)):
        for occ, vir in bf.excitations(space, rank):
            sg, d = bf._apply(bf._exc_ops(space, occ, vir), space.ref)
            Lv[space.index[d]] += sg * l[occ + vir]
    got = C.l3(H, l1, l2)
    worst = 0.0
    for occ, vir in bf.excitations(space, 3):
        sg, d = bf._apply(bf._exc_ops(space, occ, vir), space.ref)
        want = sg * (Lv @ Hb[:, space.index[d]])
        worst = max(worst, abs(got[occ + vir] - want))
    assert worst < 1e-12
    assert np.abs(got).max() > 1e-3


def _closed_shell_d3(start=None):
    n, oa = 7, 3
    s, space, hso, vso, f, Hm, amps, ec = _ccsd_state(n, oa, oa)
    no = space.nocc_so
    sm = C.SpinMaps(n, oa, oa)
    return C.d3_arrays(None, vso[:no, :no, no:, no:], amps[2], sm, d3d1_start=start), amps, sm, vso


def test_d3_spin_symmetry_and_loop_start_quirk():
    full, amps, sm, vso = _closed_shell_d3(start=0)
    np.testing.assert_allclose(full['d3d1'], full['d3a1'], atol=1e-15, rtol=0)
    np.testing.assert_allclose(full['d3d2'], full['d3a2'], atol=1e-15, rtol=0)
    np.testing.assert_allclose(full['d3c1'].transpose(0, 2, 1), full['d3b1'], atol=1e-15, rtol=0)
    np.testing.assert_allclose(full['d3c2'].transpose(1, 0, 2), full['d3b2'], atol=1e-15, rtol=0)
    quirk, _, _, _ = _closed_shell_d3()
    for k in full:
        if k != 'd3d1':
            np.testing.assert_array_equal(quirk[k], full[k])
    # closed shell: the loop starts at the second beta virtual; the dropped term is e = 0
    no = 2 * sm.oa
    t2 = amps[2]
    drop = np.zeros_like(full['d3d1'])
    for a in range(sm.vb):
        for j in range(sm.ob):
            for i in range(sm.ob):
                J, I = sm.occ(1, j), sm.occ(1, i)
                E, A = sm.vir(1, 0), sm.vir(1, a)
                drop[a, j, i] = vso[J, I, no + E, no + A] * t2[I, J, A, E]
    np.testing.assert_allclose(full['d3d1'] - quirk['d3d1'], drop, atol=1e-15, rtol=0)
    assert np.abs(drop).max() > 1e-6


def test_denominators_at_zero_amplitudes_are_hamiltonian_diagonal():
    n, oa, ob = 6, 3, 2
    s = synth.make_system(n, oa, 20, 9)
    space = bf.DetSpace(n, oa, ob)
    no, nv = space.nocc_so, space.nvir_so
    hso, vso = C.so_integrals(space, s.h, s.eri)
    fso = C.fock(hso, vso, no)
    ht, gt = C.dress(hso, vso, np.zeros((no, nv)))
    H = C.Hbar(ht, gt, np.zeros((no, no, nv, nv)), no)
    Hm = bf.hamiltonian(space, s.h, s.eri)
    e0 = Hm[space.index[space.ref], space.index[space.ref]]
    sm = C.SpinMaps(n, oa, ob)
    spin_of = lambda I, nocc_a: (0, I) if I < nocc_a else (1, I - nocc_a)
    checked = 0
    for occ, vir in bf.excitations(space, 3):
        so = [spin_of(I, sm.oa) for I in occ]
        sv = [spin_of(A, sm.va) for A in vir]
        den = C.tuple_denominators(H, fso, sm, [x for _, x in so], [x for _, x in sv],
                                   [s for s, _ in so], [s for s, _ in sv])
        _, d = bf._apply(bf._exc_ops(space, occ, vir), space.ref)
        diag = Hm[space.index[d], space.index[d]] - e0
        assert abs(-den[2] - diag) < 1e-12
        assert abs(den[0] - den[1]) < 1e-14 and den[3] == den[2]
        checked += 1
    assert checked > 50


def test_correction_closed_shell_spin_symmetry_and_full_p_space():
    from oracle import active as A
    n, oa = 7, 3
    s, space, hso, vso, fso, Hm, amps, ec = _ccsd_state(n, oa, oa)
    no, nv = space.nocc_so, space.nvir_so
    ht, gt = C.dress(hso, vso, amps[1])
    H = C.Hbar(ht, gt, amps[2], no)
    l1, l2, ok, _ = C.solve_left(space, H, amps[2], ec, tol=1e-12, max_iter=200)
    assert ok
    l3 = C.l3(H, l1, l2)
    R, _ = bf.residuals(space, Hm, {1: amps[1], 2: amps[2]}, [3])
    sm = C.SpinMaps(n, oa, oa)
    d3 = C.d3_arrays(H, vso[:no, :no, no:, no:], amps[2], sm, d3d1_start=0)
    act = A.ActiveSpace(0, oa, oa, n, 1, 1)
    cases = C.correction(H, l3, R[3], fso, d3, sm, act.m1, act.m2, by_case=True)
    np.testing.assert_allclose(cases[0], cases[3], atol=1e-14, rtol=0)
    np.testing.assert_allclose(cases[1], cases[2], atol=1e-14, rtol=0)
    assert np.abs(cases).max() > 1e-7
    full = A.ActiveSpace(0, oa, oa, n, oa, n - oa)
    assert C.correction(H, l3, R[3], fso, d3, sm, full.m1, full.m2) == (0.0, 0.0, 0.0, 0.0)


def _random_so(space, rng, rank, scale):
    from oracle.ccsdt_so import _anti
    No, Nv = space.nocc_so, space.nvir_so
    x = scale * rng.standard_normal((No,) * rank + (Nv,) * rank)
    x = _anti(x, [list(range(rank)), list(range(rank, 2 * rank))])
    return bf.blocks_to_so(space, rank, bf.so_to_blocks(space, rank, x))


@pytest.mark.parametrize("n,oa,ob", [(6, 3, 3), (6, 3, 2)])
def test_dd_is_brute_force_triples_diagonal(n, oa, ob):
    """With T1 = 0 the normal-ordered H-bar has no rank-0 piece beyond three
    bodies, so crcc23's D_D (one-, two- and the d3 three-body diagonal,
    src/mm_correct.f90:138-152,346-481, loop-start quirk off) must equal
    -(<mu|Hbar|mu> - <0|Hbar|0>) for every triple of every spin case; D_C
    (no d3) must not.  A wrong index pairing inside cal_d3 fails this."""
    s = synth.make_system(n, oa, 20, 5)
    space = bf.DetSpace(n, oa, ob)
    no, nv = space.nocc_so, space.nvir_so
    hso, vso = C.so_integrals(space, s.h, s.eri)
    Hm = bf.hamiltonian(space, s.h, s.eri)
    rng = np.random.default_rng(4)
    t1 = np.zeros((no, nv))
    t2 = _random_so(space, rng, 2, 0.1)
    H = C.Hbar(*C.dress(hso, vso, t1), t2, no)
    fso = C.fock(hso, vso, no)
    Hb = _similarity(space, Hm, {1: t1, 2: t2})
    r0 = space.index[space.ref]
    sm = C.SpinMaps(n, oa, ob)
    d3 = C.d3_arrays(H, vso[:no, :no, no:, no:], t2, sm, d3d1_start=0)
    # mistakes the identity must expose: the alpha-alpha array used for the
    # alpha-beta one, and (ccq's loop-start quirk) the first beta virtual
    # dropped from d3d1's sum
    bad = dict(d3)
    bad['d3b1'] = d3['d3a1'][:, :ob, :].copy()
    quirk = C.d3_arrays(H, vso[:no, :no, no:, no:], t2, sm, d3d1_start=1)
    gap_c = gap_bad = gap_quirk = 0.0
    count = 0
    for case in range(4):
        so_, sv_ = C.CASE_SPINS[case]
        for occ, vir in C.canonical_triples(sm, case):
            I = tuple(sm.occ(a, x) for a, x in zip(so_, occ))
            A = tuple(sm.vir(a, x) for a, x in zip(sv_, vir))
            _, d = bf._apply(bf._exc_ops(space, I, A), space.ref)
            diag = Hb[space.index[d], space.index[d]] - Hb[r0, r0]
            den = C.tuple_denominators(H, fso, sm, occ, vir, so_, sv_,
                                       C.d3_extra(d3, case, occ, vir))
            assert abs(den[3] + diag) < 1e-12, (case, occ, vir)
            gap_c = max(gap_c, abs(den[2] + diag))
            gap_bad = max(gap_bad, abs(C.d3_extra(bad, case, occ, vir)
                                       - C.d3_extra(d3, case, occ, vir)))
            gap_quirk = max(gap_quirk, abs(C.d3_extra(quirk, case, occ, vir)
                                           - C.d3_extra(d3, case, occ, vir)))
            count += 1
    assert count > 50
    assert gap_c > 1e-3 and gap_bad > 1e-3 and gap_quirk > 1e-3


@pytest.mark.parametrize("n,oa,ob", [(6, 3, 3), (6, 3, 2)])
def test_left_residual_at_nonstationary_t(n, oa, ob):
    """At any T1, T2 (not a CCSD solution: ccq solves the left equations at
    CCSDt amplitudes) and any L1, L2, the exact <0|(1+L)(Hbar-E)|mu>
    (brute force) equals the oracle's left residual -- ccq's connected terms
    plus the l1 x h_ov product of compute_l_hbar_disconnected
    (src/update_lcc.f90:106-206) -- on doubles, and on singles exceeds it by
    exactly the disconnected products sum_jb l_ij^ab R_j^b, R_j^b =
    <jb|Hbar|0>, which ccq does not include (zero at CCSD convergence)."""
    s = synth.make_system(n, oa, 20, 5)
    space = bf.DetSpace(n, oa, ob)
    no, nv = space.nocc_so, space.nvir_so
    hso, vso = C.so_integrals(space, s.h, s.eri)
    Hm = bf.hamiltonian(space, s.h, s.eri)
    rng = np.random.default_rng(8)
    t1, t2 = _random_so(space, rng, 1, 0.05), _random_so(space, rng, 2, 0.1)
    l1, l2 = _random_so(space, rng, 1, 0.1), _random_so(space, rng, 2, 0.1)
    H = C.Hbar(*C.dress(hso, vso, t1), t2, no)
    Hb = _similarity(space, Hm, {1: t1, 2: t2})
    r0 = space.index[space.ref]
    E = Hb[r0, r0]
    basis = _sd_basis(space)
    lam = np.array([l1[oc[0], vi[0]] if rk == 1 else l2[oc + vi] for rk, oc, vi, _, _ in basis])
    R1 = np.zeros((no, nv))
    for rk, oc, vi, sg, d in basis:
        if rk == 1:
            R1[oc[0], vi[0]] = sg * Hb[d, r0]
    assert np.abs(R1).max() > 1e-2
    extra = np.einsum('ijab,jb->ia', l2, R1)
    r1, r2 = C.lambda_residual(H, t2, l1, l2)
    r1_bad, _ = C.lambda_residual(H, 0 * t2, l1, l2)     # Gae / Gmi dropped
    worst_bad = 0.0
    for m, (rk, oc, vi, sm_, dm) in enumerate(basis):
        brute = Hb[r0, dm] * sm_
        for k, (_, _, _, sn, dn) in enumerate(basis):
            brute += lam[k] * sn * sm_ * (Hb[dn, dm] - (E if dn == dm else 0.0))
        if rk == 1:
            i, a = oc[0], vi[0]
            assert abs(brute - (r1[i, a] + extra[i, a])) < 1e-12
            worst_bad = max(worst_bad, abs(brute - (r1_bad[i, a] + extra[i, a])))
        else:
            assert abs(brute - r2[oc + vi]) < 1e-12
    assert worst_bad > 1e-4
