"""Pins for the learned-guess oracle (oracle/mlp.py): the BF16 rounding
against torch's float32 -> bfloat16 conversion (round to nearest even), the
gradient against central finite differences (rounding off, float64), and
the MP2 amplitude of the features against the first CCSD Jacobi step from
T = 0 (ccq's first T2A update is -V / D with D = f_aa + f_bb - f_ii - f_jj)."""

import numpy as np
import torch

from oracle import integrals as I
from oracle import mlp
from oracle import solver as S
from ccq_amd import synth


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 3, 20000), rng.normal(0, 1e-3, 2000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0])]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.bf16(x), want)


def _params(seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp.NPARAMS)
    a, b = mlp.OFF["sd"]
    p[a:b] = rng.uniform(0.5, 2.0, b - a)
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    X = rng.normal(0, 1, (64, mlp.F))
    M = rng.normal(0, 1, 64)
    Y = rng.normal(0, 0.3, 64)
    p = _params(2)
    loss, g = mlp.loss_grad(p, X, M, Y, rounding=False)
    h = 1e-6
    a0, _ = mlp.OFF["W1"]
    for k in list(range(a0, mlp.NPARAMS, 97)) + [mlp.NPARAMS - 1]:
        dp = np.zeros_like(p)
        dp[k] = h
        lp, _ = mlp.loss_grad(p + dp, X, M, Y, rounding=False)
        lm, _ = mlp.loss_grad(p - dp, X, M, Y, rounding=False)
        assert abs((lp - lm) / (2 * h) - g[k]) < 1e-7 * max(1.0, abs(g[k])), k
    # mu / sd are not trained
    assert not g[:2 * mlp.F].any()


def test_mp2_feature_is_first_jacobi_step():
    n, o = 12, 4
    s = synth.make_system(n, o, 24, 3)
    e2 = s.e2int()
    fa, fb = I.gen_fock(s.h, e2, o, o)
    vaa, vab, vbb = I.antisymmetrize(e2, 0)
    t = S.zero_amplitudes('CCSD', o, o, n - o, n - o)
    D = S.block_denominators(t, fa, fb, o, o, 0.0)
    R = S.CCSDResiduals(s.h, e2, 0, o, o)('t2aa', t)
    first = -R / D['t2aa']
    fd = np.diag(fa)
    X, m, inv = mlp.features(vaa.reshape(n, n, n, n), fd, fd, o, o, o, o)
    np.testing.assert_allclose(m, first, atol=1e-15, rtol=0)
    np.testing.assert_allclose(X[:, 0], np.abs(m).ravel() * inv, rtol=1e-15)
    assert abs(np.sqrt(np.mean((m * inv) ** 2)) - 1.0) < 1e-12
