"""Pins of the MLP curvature-closure oracle (oracle/mlp_oracle.py), CPU only.

The network has no counterpart in the reference, so its oracle is pinned against
independent arithmetic: bfloat16 rounding against hand-worked round-to-nearest-even
cases, and the backward pass against central finite differences of the FP64 loss
(a dropped activation derivative, a transposed weight or a missing 1/B would fail).
"""
import numpy as np
import pytest

import mlp_oracle as M


def test_bf16_rounding_cases():
    u = 2.0 ** -7                                      # bf16 ulp at 1
    cases = [
        (1.0, 1.0),
        (1.0 + u / 2, 1.0),                            # tie -> even mantissa (0)
        (1.0 + 1.5 * u, 1.0 + 2 * u),                  # tie -> even mantissa (2)
        (1.0 + u / 2 + 2.0 ** -20, 1.0 + u),           # just above the tie
        (1.0 + u / 2 - 2.0 ** -20, 1.0),               # just below the tie
        (-3.0 - 0.5 * 2 * u, -3.0),                    # ulp in [2, 4) is 2u; tie -> even (3 = 1.1000000b x 2)
        (-3.0 - 1.5 * 2 * u, -3.0 - 2 * 2 * u),        # tie between odd and even -> even
        (0.0, 0.0),
    ]
    for x, want in cases:
        assert M.bf16(np.array([x]))[0] == want, (x, M.bf16(np.array([x]))[0], want)


def test_bf16_is_idempotent_and_bounded():
    rng = np.random.default_rng(3)
    x = rng.standard_normal(10000) * 10.0 ** rng.integers(-5, 5, 10000)
    r = M.bf16(x)
    assert np.array_equal(M.bf16(r), r)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * (1 + 1e-12))


def _flat(p):
    return np.concatenate([p[k].ravel() for k in ("W1", "b1", "W2", "b2", "w3", "b3")])


def _unflat(p, v):
    q, o = {}, 0
    for k in ("W1", "b1", "W2", "b2", "w3", "b3"):
        n = p[k].size
        q[k] = v[o:o + n].reshape(p[k].shape).copy()
        o += n
    return q


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(11)
    p = M.init_params(5)
    for k in ("b1", "b2", "b3"):
        p[k] = rng.standard_normal(p[k].shape) * 0.1
    B = 17
    X = np.zeros((B, M.NPAD))
    X[:, :M.NIN] = rng.standard_normal((B, M.NIN))
    t = rng.standard_normal(B)
    _, g = M.loss_and_grads(p, X, t)
    v0 = _flat(p)
    gv = _flat(g)
    h = 1e-6
    idx = rng.choice(len(v0), 60, replace=False)
    # every bias and the output weights, plus a random sample of the matrices
    for q in idx:
        vp, vm = v0.copy(), v0.copy()
        vp[q] += h
        vm[q] -= h
        lp, _ = M.loss_and_grads(_unflat(p, vp), X, t)
        lm, _ = M.loss_and_grads(_unflat(p, vm), X, t)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gv[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, gv[q])


def test_padding_columns_get_zero_gradient():
    rng = np.random.default_rng(2)
    p = M.init_params(1)
    X = np.zeros((9, M.NPAD))
    X[:, :M.NIN] = rng.standard_normal((9, M.NIN))
    _, g = M.loss_and_grads(p, X, rng.standard_normal(9))
    assert np.all(g["W1"][:, M.NIN:] == 0.0)


def test_bf16_gradients_close_to_fp64():
    rng = np.random.default_rng(4)
    p = M.init_params(2)
    X = np.zeros((64, M.NPAD))
    X[:, :M.NIN] = rng.standard_normal((64, M.NIN))
    t = rng.standard_normal(64)
    _, g = M.loss_and_grads(p, X, t)
    _, gb = M.loss_and_grads(p, X, t, bf=True)
    for k in g:
        assert np.max(np.abs(gb[k] - g[k])) <= 3e-2 * max(np.max(np.abs(g[k])), 1e-12), k
