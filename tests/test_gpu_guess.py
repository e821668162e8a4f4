"""GPU tests of the learned T2 starting guess (ccq_amd/csrc/guess.hip) against
the CPU oracle (oracle/mlp.py), and what the shipped network gains.

Bar: MP2 guess == V / D to 1e-14; training data features / targets to FP32
roundoff; the GPU loss and gradient of the training kernels == the oracle's
(FP32 accumulation, BF16 rounding points identical: 1e-3 of the largest
gradient entry, loss 1e-4 relative); learned inference == the oracle's
forward for 99.9% of the tuples to 1e-5 |m| (a value within FP32 roundoff of
a BF16 rounding midpoint may round the other way); training lowers the loss;
with ccq_amd/guess_weights.npz the learned guess reaches |E - E_conv| < 1e-7
in fewer iterations than the MP2 guess on held-out systems, and every guess
converges to the same energy (1e-9 Eh).
"""

import os
import sys

import numpy as np
import pytest

from oracle import mlp
from ccq_amd import engine as E
from ccq_amd import guess as GS
from ccq_amd import solve as SV

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_blocks(eng_spec):
    """Oracle features per T2 block of a training-style system."""
    from ccq_amd import synth
    from ccq_amd.inputs import build_integrals
    n, o, L, seed, sigma = eng_spec
    s = synth.make_system(n, o, L, seed, sigma=sigma)
    ints = build_integrals(s.h, s.e2int(), s.enuc, 0, o, o)
    fd = np.diag(ints.f_a)
    out = []
    for V in (ints.v_aa, ints.v_ab, ints.v_bb):
        out.append(mlp.features(np.asarray(V).reshape(n, n, n, n), fd, fd, o, o, o, o))
    return out


SPEC = (14, 4, 28, 77, 0.05)


def _spec():
    from ccq_amd import synth
    n, o, L, seed, _ = SPEC
    return (n, o, L, seed, synth.regime_sigma(n, o, L))


def test_mp2_guess_is_closed_form():
    _say("case guess mp2")
    spec = _spec()
    eng = GS.make_engine(*spec)
    eng.guess("mp2")
    blocks = E.split_tvec(eng.get_t(), spec[1], spec[1], spec[0] - spec[1], spec[0] - spec[1])
    want = _oracle_blocks(spec)
    for name, (X, m, inv) in zip(("t2aa", "t2ab", "t2bb"), want):
        np.testing.assert_allclose(blocks[name], m, atol=1e-14, rtol=0)
    assert not blocks["t1a"].any() and not blocks["t1b"].any()
    eng.guess("zero")
    assert not eng.get_t().any()
    eng.close()


def test_guess_data_matches_oracle():
    _say("case guess data")
    spec = _spec()
    eng = GS.make_engine(*spec)
    r = SV.jacobi_iter(eng, 1e-10, 100, 5, out=None)
    assert r.converged
    X, M, Y = eng.guess_data()
    n, o = spec[0], spec[1]
    blocks = E.split_tvec(eng.get_t(), o, o, n - o, n - o)
    Xw, Mw, Yw = [], [], []
    for name, (Xo, m, inv) in zip(("t2aa", "t2ab", "t2bb"), _oracle_blocks(spec)):
        Xw.append(Xo)
        Mw.append((m * inv).ravel())
        Yw.append(((blocks[name] - m) * inv).ravel())
    for got, want in ((X, np.concatenate(Xw)), (M, np.concatenate(Mw)), (Y, np.concatenate(Yw))):
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6 * np.abs(want).max())
    eng.close()


def _data(seed, n=3000):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, 5)).astype(np.float32)
    X[:, 0] = np.abs(X[:, 0])
    M = rng.normal(0, 1, n).astype(np.float32)
    Y = (0.3 * M * np.tanh(X[:, 0]) + 0.05 * rng.normal(0, 1, n)).astype(np.float32)
    return X, M, Y


def test_training_kernel_gradient_matches_oracle():
    _say("case mlp gradient")
    X, M, Y = _data(3)
    p = GS.init_params(X, seed=4)
    loss, g = GS.grad(X, M, Y, p)
    lw, gw = mlp.loss_grad(p.astype(np.float64), X, M, Y, rounding=True)
    assert abs(loss - lw) < 1e-4 * lw
    assert np.abs(g - gw).max() < 1e-3 * np.abs(gw).max()
    assert not g[:10].any()


def test_learned_inference_matches_oracle():
    _say("case learned inference")
    spec = _spec()
    X, M, Y = _data(5)
    p = GS.init_params(X, seed=6)
    p, _ = GS.train(X, M, Y, p, steps=50, batch=512)
    eng = GS.make_engine(*spec)
    eng.guess("learned", p)
    n, o = spec[0], spec[1]
    blocks = E.split_tvec(eng.get_t(), o, o, n - o, n - o)
    errs, scale = [], []
    for name, (Xo, m, inv) in zip(("t2aa", "t2ab", "t2bb"), _oracle_blocks(spec)):
        r, _ = mlp.forward(p.astype(np.float64), Xo.astype(np.float32).astype(np.float64))
        want = m * (1.0 + r.reshape(m.shape))
        errs.append(np.abs(blocks[name] - want).ravel())
        scale.append(np.abs(m).ravel())
    err, sc = np.concatenate(errs), np.concatenate(scale)
    ok = err <= 1e-5 * sc + 1e-15
    assert ok.mean() > 0.999, ok.mean()
    assert err.max() < 5e-2 * sc.max()
    # antisymmetric blocks stay antisymmetric
    t2aa = blocks["t2aa"]
    assert np.abs(t2aa + t2aa.transpose(1, 0, 2, 3)).max() < 1e-15
    eng.close()


def test_training_lowers_the_loss():
    _say("case mlp training")
    from ccq_amd import synth
    systems = [(n, o, 2 * n, 500 + n, synth.regime_sigma(n, o, 2 * n)) for n, o in ((12, 3), (14, 4))]
    X, M, Y = GS.collect(systems)
    p0 = GS.init_params(X)
    l0, _ = GS.grad(X, M, Y, p0)
    p, _ = GS.train(X, M, Y, p0, steps=300, batch=1024)
    l1, _ = GS.grad(X, M, Y, p)
    assert l1 < 0.5 * l0, (l0, l1)


def _iters_to(eng, mode, params, e_conv):
    eng.guess(mode, params)
    r = SV.jacobi_iter(eng, 1e-11, 100, 5, out=None)
    assert r.converged
    hit = next(h[0] for h in r.history if abs(h[1] - e_conv) < 1e-7)
    return hit, r.en_cor


@pytest.mark.skipif(not os.path.exists(GS.WEIGHTS), reason="ccq_amd/guess_weights.npz not trained")
@pytest.mark.parametrize("n,o,seed", [(28, 5, 7), (40, 6, 9)])
def test_learned_guess_gains_over_mp2(n, o, seed):
    _say(f"case learned gain n={n} o={o}")
    from ccq_amd import synth
    params = GS.load_weights()
    eng = GS.make_engine(n, o, 2 * n, seed, synth.regime_sigma(n, o, 2 * n))
    ref = SV.jacobi_iter(eng, 1e-11, 100, 5, out=None)
    e_conv = ref.en_cor
    it_mp2, e_mp2 = _iters_to(eng, "mp2", None, e_conv)
    it_ml, e_ml = _iters_to(eng, "learned", params, e_conv)
    _say(f"  iterations to 1e-7: zero {next(h[0] for h in ref.history if abs(h[1] - e_conv) < 1e-7)}, "
         f"mp2 {it_mp2}, learned {it_ml}")
    assert it_ml < it_mp2
    assert abs(e_ml - e_conv) < 1e-9 and abs(e_mp2 - e_conv) < 1e-9
    eng.close()
