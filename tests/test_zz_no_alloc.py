"""ccq_hip_mlp_grad_noalloc: the gradient entry point that allocates no
device memory (static per-GPU scratch) gives the results of
ccq_hip_mlp_grad, bit for bit, on two calls in a row (the second with fewer
samples, so stale scratch rows beyond n must not leak in), and refuses a
sample count beyond its limit without touching the GPU."""

import ctypes
import os
import re
import sys

import numpy as np
import pytest

from ccq_amd import engine as E
from ccq_amd import guess as GS

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _limit():
    text = open(os.path.join(ROOT, "include", "ccq_hip.h")).read()
    return int(re.search(r"#define CCQ_HIP_MLP_NOALLOC_MAX_SAMPLES (\d+)", text).group(1))


def _data(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, 5)).astype(np.float32)
    M = rng.normal(0, 1, n).astype(np.float32)
    Y = (0.3 * M * np.tanh(X[:, 0]) + 0.05 * rng.normal(0, 1, n)).astype(np.float32)
    return X, M, Y


def test_noalloc_gradient_equals_allocating_gradient():
    nmax = _limit()
    for call, (seed, n) in enumerate(((11, nmax), (12, 1000))):
        _say(f"case noalloc gradient: call {call + 1}, n = {n}")
        X, M, Y = _data(seed, n)
        p = GS.init_params(X, seed=seed)
        l_ref, g_ref = GS.grad(X, M, Y, p)
        l_na, g_na = GS.grad(X, M, Y, p, noalloc=True)
        assert l_na == l_ref
        assert np.array_equal(g_na, g_ref)
        assert np.abs(g_ref).max() > 0


def test_noalloc_gradient_rejects_too_many_samples():
    _say("case noalloc gradient: size limit")
    n = _limit() + 1
    X, M, Y = _data(13, n)
    p = GS.init_params(X, seed=1)
    with pytest.raises(E.EngineError, match="CCQ_HIP_MLP_NOALLOC_MAX_SAMPLES"):
        GS.grad(X, M, Y, p, noalloc=True)
    lib = E.load_library()
    g = np.zeros(p.size, dtype=np.float32)
    rc = lib.ccq_hip_mlp_grad_noalloc(None, None, None, 0, p.ctypes.data_as(ctypes.c_void_p),
                                      p.size, g.ctypes.data_as(ctypes.c_void_p), None)
    assert rc == 2
