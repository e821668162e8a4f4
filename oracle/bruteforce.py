"""Brute-force determinant-space coupled cluster (oracle; tiny systems only).

Independent of any factorisation of the CC equations: the residual of a
spin-orbital cluster operator T is computed from its definition

    R_mu = <Phi_mu| exp(-T) H exp(T) |Phi>,

with H, T and the determinants represented explicitly in the N-electron,
fixed-M_s determinant space.  This pins every residual block of CCSD, CCSDT
and CCSDTQ (ccq's src/cc/t{1..4}{A..E}_update.f90 compute exactly these
projections, factorised) at any amplitude values, and the FCI energy pins
CCSDTQ for 4 electrons and CCSD for 2 electrons.

Conventions (shared with the block layout used everywhere in this repo):
  * spin-orbital tensors t_k have k occupied then k virtual indices,
    fully antisymmetric within each group;
  * occupied spin orbitals: alpha 0..oa-1 then beta 0..ob-1;
    virtual spin orbitals: alpha 0..va-1 then beta 0..vb-1;
  * |Phi_{I1..Ik}^{A1..Ak}> = a+_{A1}..a+_{Ak} a_{Ik}..a_{I1} |Phi>, the
    ordering ccq uses (|Phi_ij^ab> = a+ b+ j i |Phi>).
Spin blocks: block ``nb`` of rank k has the last nb occupied and the last nb
virtual indices beta (alpha indices first), e.g. rank 2: aa, ab, bb.
"""

from __future__ import annotations

import itertools
import math

import numpy as np
import scipy.sparse as sp


def _parity(perm) -> int:
    perm = list(perm)
    sign = 1
    seen = [False] * len(perm)
    for i in range(len(perm)):
        if seen[i]:
            continue
        j = i
        cyc = 0
        while not seen[j]:
            seen[j] = True
            j = perm[j]
            cyc += 1
        if cyc % 2 == 0:
            sign = -sign
    return sign


class DetSpace:
    """Determinants with fixed (N_alpha, N_beta) over n spatial orbitals."""

    def __init__(self, norb: int, oa: int, ob: int):
        self.n, self.oa, self.ob = norb, oa, ob
        self.va, self.vb = norb - oa, norb - ob
        alpha = [sum(1 << p for p in c) for c in itertools.combinations(range(norb), oa)]
        beta = [sum(1 << (norb + p) for p in c) for c in itertools.combinations(range(norb), ob)]
        self.dets = [a | b for a in alpha for b in beta]
        self.index = {d: i for i, d in enumerate(self.dets)}
        self.ref = sum(1 << p for p in range(oa)) | sum(1 << (norb + p) for p in range(ob))

    # spin-orbital tensor index -> determinant bit
    def occ_bit(self, I: int) -> int:
        return I if I < self.oa else self.n + (I - self.oa)

    def vir_bit(self, A: int) -> int:
        return self.oa + A if A < self.va else self.n + self.ob + (A - self.va)

    @property
    def nocc_so(self) -> int:
        return self.oa + self.ob

    @property
    def nvir_so(self) -> int:
        return self.va + self.vb

    def spin_of_occ(self, I: int) -> int:
        return 0 if I < self.oa else 1

    def spin_of_vir(self, A: int) -> int:
        return 0 if A < self.va else 1


def _apply(ops, det):
    """Apply a string of (bit, is_creator) right-to-left; return (sign, det) or None."""
    sign = 1
    for bit, cre in reversed(ops):
        m = 1 << bit
        occ = bool(det & m)
        if cre == occ:
            return None
        if bin(det & (m - 1)).count("1") & 1:
            sign = -sign
        det = det | m if cre else det & ~m
    return sign, det


def spin_orbital_integrals(h: np.ndarray, eri: np.ndarray):
    """Spin-orbital h_PQ and <PQ||RS> over 2n orbitals (bit order: alpha p, beta n+p)."""
    n = h.shape[0]
    N = 2 * n
    spin = np.array([0] * n + [1] * n)
    sp_idx = np.array(list(range(n)) + list(range(n)))
    hso = np.where(spin[:, None] == spin[None, :], h[np.ix_(sp_idx, sp_idx)], 0.0)
    # <PQ|RS> = (PR|QS)
    g = eri[np.ix_(sp_idx, sp_idx, sp_idx, sp_idx)].transpose(0, 2, 1, 3)
    same_pr = spin[:, None] == spin[None, :]
    mask = same_pr[:, None, :, None] & same_pr[None, :, None, :]
    g = np.where(mask, g, 0.0)
    vso = g - g.transpose(0, 1, 3, 2)
    assert vso.shape == (N, N, N, N)
    return hso, vso


def hamiltonian(space: DetSpace, h: np.ndarray, eri: np.ndarray) -> np.ndarray:
    """Dense H in the determinant basis (electronic part only)."""
    hso, vso = spin_orbital_integrals(h, eri)
    N = 2 * space.n
    dim = len(space.dets)
    H = np.zeros((dim, dim))
    for col, det in enumerate(space.dets):
        occ = [p for p in range(N) if det >> p & 1]
        emp = [p for p in range(N) if not det >> p & 1]
        # one-body
        for q in occ:
            for p in [q] + emp:
                val = hso[p, q]
                if val == 0.0:
                    continue
                res = _apply([(p, True), (q, False)], det)
                if res is None:
                    continue
                s, d2 = res
                H[space.index[d2], col] += s * val
        # two-body: sum_{P<Q, R<S} <PQ||RS> a+P a+Q aS aR
        for r, s_ in itertools.combinations(occ, 2):
            avail = sorted(emp + [r, s_])
            for p, q in itertools.combinations(avail, 2):
                val = vso[p, q, r, s_]
                if val == 0.0:
                    continue
                res = _apply([(p, True), (q, True), (s_, False), (r, False)], det)
                if res is None:
                    continue
                sg, d2 = res
                H[space.index[d2], col] += sg * val
    return H


def excitations(space: DetSpace, rank: int):
    """Unique rank-k excitations: (occ tuple, vir tuple) in increasing order."""
    for occ in itertools.combinations(range(space.nocc_so), rank):
        so = sum(space.spin_of_occ(I) for I in occ)
        for vir in itertools.combinations(range(space.nvir_so), rank):
            if sum(space.spin_of_vir(A) for A in vir) != so:
                continue
            yield occ, vir


def _exc_ops(space: DetSpace, occ, vir):
    return ([(space.vir_bit(A), True) for A in vir] +
            [(space.occ_bit(I), False) for I in reversed(occ)])


def t_operator(space: DetSpace, amps: dict) -> sp.csr_matrix:
    """Sparse matrix of T = sum_k sum_{unique} t_k[occ, vir] E_{occ,vir}."""
    rows, cols, vals = [], [], []
    for rank, t in amps.items():
        for occ, vir in excitations(space, rank):
            amp = t[occ + vir]
            if amp == 0.0:
                continue
            ops = _exc_ops(space, occ, vir)
            for col, det in enumerate(space.dets):
                res = _apply(ops, det)
                if res is None:
                    continue
                s, d2 = res
                rows.append(space.index[d2])
                cols.append(col)
                vals.append(s * amp)
    dim = len(space.dets)
    return sp.csr_matrix((vals, (rows, cols)), shape=(dim, dim))


def _expm_apply(T, v, sign=1.0):
    out = v.copy()
    term = v.copy()
    for k in range(1, 64):
        term = (sign / k) * (T @ term)
        if not np.any(term):
            break
        out += term
    return out


def antisym_fill(rank: int, unique: dict, nocc: int, nvir: int) -> np.ndarray:
    """Full antisymmetric tensor from values on increasing (occ, vir) tuples."""
    t = np.zeros((nocc,) * rank + (nvir,) * rank)
    perms = [(p, _parity(p)) for p in itertools.permutations(range(rank))]
    for (occ, vir), val in unique.items():
        for po, so in perms:
            o2 = tuple(occ[i] for i in po)
            for pv, sv in perms:
                v2 = tuple(vir[i] for i in pv)
                t[o2 + v2] = so * sv * val
    return t


def residuals(space: DetSpace, H: np.ndarray, amps: dict, ranks) -> tuple[dict, float]:
    """Return ({rank: R_k full antisymmetric}, <Phi|exp(-T) H exp(T)|Phi>)."""
    T = t_operator(space, amps)
    phi = np.zeros(len(space.dets))
    phi[space.index[space.ref]] = 1.0
    v = _expm_apply(T, phi, +1.0)
    v = H @ v
    v = _expm_apply(T, v, -1.0)
    out = {}
    for rank in ranks:
        uniq = {}
        for occ, vir in excitations(space, rank):
            res = _apply(_exc_ops(space, occ, vir), space.ref)
            s, d = res
            uniq[(occ, vir)] = s * v[space.index[d]]
        out[rank] = antisym_fill(rank, uniq, space.nocc_so, space.nvir_so)
    return out, float(v[space.index[space.ref]])


def fci_energy(H: np.ndarray) -> float:
    return float(np.linalg.eigvalsh(H)[0])


# --------------------------------------------------------------------------
# spin blocks <-> spin-orbital tensors


def block_shape(rank: int, nb: int, oa: int, ob: int, va: int, vb: int):
    na = rank - nb
    return (oa,) * na + (ob,) * nb + (va,) * na + (vb,) * nb


def _spin_pattern_map(rank, nb, space: DetSpace):
    """Yield (occ_so_index_lists, vir_so_index_lists) offsets for a spin block."""
    na = rank - nb
    o_off = [0] * na + [space.oa] * nb
    v_off = [0] * na + [space.va] * nb
    return o_off, v_off


def blocks_to_so(space: DetSpace, rank: int, blocks: dict) -> np.ndarray:
    """Assemble a full spin-orbital tensor from spin blocks {nb: array}."""
    No, Nv = space.nocc_so, space.nvir_so
    t = np.zeros((No,) * rank + (Nv,) * rank)
    perms = [(p, _parity(p)) for p in itertools.permutations(range(rank))]
    for nb, blk in blocks.items():
        o_off, v_off = _spin_pattern_map(rank, nb, space)
        # place the canonical block, then all index permutations with parity
        ix = [np.arange(blk.shape[k]) + o_off[k] for k in range(rank)] + \
             [np.arange(blk.shape[rank + k]) + v_off[k] for k in range(rank)]
        for po, so in perms:
            for pv, sv in perms:
                axes_o = [ix[po[k]] for k in range(rank)]
                axes_v = [ix[rank + pv[k]] for k in range(rank)]
                tb = blk.transpose(list(po) + [rank + pv[k] for k in range(rank)])
                t[np.ix_(*axes_o, *axes_v)] = so * sv * tb
    return t


def so_to_blocks(space: DetSpace, rank: int, t: np.ndarray) -> dict:
    out = {}
    for nb in range(rank + 1):
        o_off, v_off = _spin_pattern_map(rank, nb, space)
        shp = block_shape(rank, nb, space.oa, space.ob, space.va, space.vb)
        ix = [np.arange(shp[k]) + o_off[k] for k in range(rank)] + \
             [np.arange(shp[rank + k]) + v_off[k] for k in range(rank)]
        out[nb] = t[np.ix_(*ix)].copy()
    return out


def solve_cc(space: DetSpace, H: np.ndarray, max_rank: int, fock_diag_o, fock_diag_v,
             tol: float = 1e-12, maxiter: int = 500):
    """Converge spin-orbital CC (ranks 1..max_rank) by Jacobi + DIIS.

    Returns (amplitudes {rank: tensor}, correlation energy).  The update is
    t -= R / D with D the orbital-energy difference, which has the same fixed
    point R = 0 whatever the step; DIIS accelerates it.
    """
    No, Nv = space.nocc_so, space.nvir_so
    ranks = list(range(1, max_rank + 1))
    amps = {k: np.zeros((No,) * k + (Nv,) * k) for k in ranks}
    Dk = {}
    for k in ranks:
        d = np.zeros((No,) * k + (Nv,) * k)
        for idx in np.ndindex(*d.shape):
            d[idx] = sum(fock_diag_v[a] for a in idx[k:]) - sum(fock_diag_o[i] for i in idx[:k])
        Dk[k] = d
    phi = np.zeros(len(space.dets))
    phi[space.index[space.ref]] = 1.0
    eref = float(phi @ H @ phi)
    hist_t, hist_e = [], []
    ecorr = 0.0
    for it in range(maxiter):
        R, e = residuals(space, H, amps, ranks)
        ecorr = e - eref
        err = max(np.abs(R[k]).max() for k in ranks)
        if err < tol:
            break
        new = {k: amps[k] - R[k] / Dk[k] for k in ranks}
        vec = np.concatenate([new[k].ravel() for k in ranks])
        dv = np.concatenate([(new[k] - amps[k]).ravel() for k in ranks])
        hist_t.append(vec)
        hist_e.append(dv)
        hist_t, hist_e = hist_t[-8:], hist_e[-8:]
        if len(hist_t) >= 3:
            m = len(hist_t)
            B = -np.ones((m + 1, m + 1))
            B[m, m] = 0.0
            for a in range(m):
                for b in range(m):
                    B[a, b] = hist_e[a] @ hist_e[b]
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            c = np.linalg.lstsq(B, rhs, rcond=None)[0][:m]
            vec = sum(ci * ti for ci, ti in zip(c, hist_t))
        off = 0
        for k in ranks:
            sz = amps[k].size
            amps[k] = vec[off:off + sz].reshape(amps[k].shape)
            off += sz
    return amps, ecorr


def math_comb(n: int, k: int) -> int:
    return math.comb(n, k)
