"""CPU restatement of the learned-guess network (oracle; TEST INFRASTRUCTURE ONLY).

The engine's guess (ccq_amd/csrc/guess.hip) has no counterpart in ccq (ccq
starts from T = 0, src/calc_driver.f90:210-211); this module restates its
definition in numpy float64 so the GPU kernels can be checked:

* features of a T2 tuple (i, j, a, b) of a spin block with V = <ij||ab>
  (<ij|ab> for alpha-beta), D = f_ii + f_jj - f_aa - f_bb, m = V / D:
  x = (|m| / sm, f_ii + f_jj, f_aa + f_bb, V_ijij, V_abab), sm = rms(m);
* network x -> (x - mu) / sd -> 64 tanh -> 64 tanh -> 1, with the weights and
  every layer input rounded to BF16 (nearest, ties to even); guess
  t = m (1 + r);
* loss = mean (M r - Y)^2 with M = m / sm, Y = (t_conv - m) / sm; gradients
  with respect to the master weights through the rounded values
  (straight-through), tanh' taken as 1 - h^2 of the rounded activation.

Pinned by tests/test_oracle_mlp.py (finite differences with the rounding
switched off, MP2 closed form of the features).
"""

from __future__ import annotations

import numpy as np

F, H = 5, 64
NPARAMS = 2 * F + F * H + H + H * H + H + H + 1


def offsets():
    o = {}
    k = 0
    for name, size in (("mu", F), ("sd", F), ("W1", F * H), ("b1", H), ("W2", H * H), ("b2", H),
                       ("W3", H), ("b3", 1)):
        o[name] = (k, k + size)
        k += size
    return o


OFF = offsets()


def unpack(p):
    g = {k: np.asarray(p[a:b], dtype=np.float64) for k, (a, b) in OFF.items()}
    g["W1"] = g["W1"].reshape(F, H)
    g["W2"] = g["W2"].reshape(H, H)
    return g


def bf16(x):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def features(V4, fdiag1, fdiag2, o1, o2, va0, vb0):
    """Features, m and sm of one spin block from full-range integrals V4[p,q,r,s]
    (occupied first per spin) and [occ | vir] Fock diagonals of the two spins."""
    v1 = len(fdiag1) - o1
    v2 = len(fdiag2) - o2
    i = np.arange(o1)[:, None, None, None]
    j = np.arange(o2)[None, :, None, None]
    a = np.arange(v1)[None, None, :, None]
    b = np.arange(v2)[None, None, None, :]
    V = V4[i, j, va0 + a, vb0 + b]
    so = fdiag1[i] + fdiag2[j]
    sv = fdiag1[o1 + a] + fdiag2[o2 + b]
    m = V / (so - sv)
    sm = np.sqrt(np.mean(m ** 2)) if m.size else 0.0
    inv = 1.0 / sm if sm > 0 else 0.0
    X = np.stack(np.broadcast_arrays(np.abs(m) * inv, so, sv, V4[i, j, i, j],
                                     V4[va0 + a, vb0 + b, va0 + a, vb0 + b]), -1)
    return X.reshape(-1, F), m, inv


def forward(p, X, rounding=True):
    """r and the kept activations (rounded input, h1, h2)."""
    g = unpack(p)
    rnd = bf16 if rounding else (lambda z: np.asarray(z, dtype=np.float64))
    W1, W2, W3 = rnd(g["W1"]), rnd(g["W2"]), rnd(g["W3"])
    x = rnd((X.astype(np.float64) - g["mu"]) / g["sd"])
    h1 = rnd(np.tanh(x @ W1 + g["b1"]))
    h2 = rnd(np.tanh(h1 @ W2 + g["b2"]))
    r = h2 @ W3 + g["b3"][0]
    return r, (x, h1, h2, W2, W3)


def loss_grad(p, X, M, Y, rounding=True):
    """Mean loss and its gradient (mu / sd entries zero) over all samples."""
    r, (x, h1, h2, W2, W3) = forward(p, X, rounding)
    n = len(r)
    e = M * r - Y
    g = 2.0 * e * M / n
    gh2 = g[:, None] * W3[None, :] * (1.0 - h2 ** 2)
    gh1 = (gh2 @ W2.T) * (1.0 - h1 ** 2)
    grad = np.zeros(NPARAMS)
    put = lambda name, val: grad.__setitem__(slice(*OFF[name]), np.ravel(val))
    put("W1", x.T @ gh1)
    put("b1", gh1.sum(0))
    put("W2", h1.T @ gh2)
    put("b2", gh2.sum(0))
    put("W3", h2.T @ g)
    put("b3", [g.sum()])
    return float(np.mean(e ** 2)), grad
