"""Pins the FP64 MLP oracle (oracle/mlp_oracle.py) with checks that do not
restate it: finite differences, closed forms, special values from the
reference's source lines, and the C oracle's PCG32."""
import ctypes as C

import numpy as np
import pytest

import oracle
from oracle import mlp_oracle as M


def _params(seed, scale=0.2):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, scale, M.N_PARAMS)
    p[5 * 4096 + 3 * 64:] = 0.0
    return p


def _verts(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.uniform(0, 1, (n, 16))
    v[:, 0:3] = rng.uniform(-6, 2, (n, 3))
    return v


def test_gradient_matches_finite_differences():
    p = _params(1)
    x = M.encode(_verts(32, 2), (-6, -4, -10), (6, 6, 2))
    tgt = np.random.default_rng(3).uniform(0, 2, (32, 3))
    _, g = M.loss_and_grad(p, x, tgt)
    rng = np.random.default_rng(4)
    idx = np.concatenate([rng.choice(5 * 4096, 40, replace=False), 5 * 4096 + rng.choice(3 * 64, 10, replace=False)])
    h = 1e-6
    for i in idx:
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (M.loss_only(pp, x, tgt, p) - M.loss_only(pm, x, tgt, p)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
    # padding rows of the output layer never receive gradient
    assert not g[5 * 4096 + 3 * 64:].any()


def test_forward_is_plain_relu_mlp():
    p = _params(5)
    x = np.random.default_rng(6).normal(size=(8, 64))
    pre, ins = M.forward(p, x)
    Ws = M.split(p)
    h = x
    for l in range(5):
        h = np.maximum(h @ Ws[l].T, 0)
    np.testing.assert_allclose(pre[5], h @ Ws[5].T, rtol=1e-12)


def test_bf16_rounding_is_round_to_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0])
    r = M.round_to(v, "bf16")
    np.testing.assert_array_equal(r, [1.0, 1.0, 1.0 + 2 ** -7, -2.5, 0.0])


def test_batch_indices_match_c_pcg32():
    L = oracle.lib()
    for seed in (0, 1, 7, 12345):
        idx = M.batch_indices(seed, 64, 1000)
        for i in (0, 1, 17, 63):
            s = C.c_uint32(seed ^ i)
            assert idx[i] == L.orc_pcg32(C.byref(s)) % 1000


def test_encoding_closed_forms():
    v = np.zeros((3, 16))
    v[0, 0:3] = (-6, -4, -10)     # scene minimum -> all sin features 0
    v[1, 0:3] = (0.0, 1.0, -4.0)  # centre: x = 0.5
    v[:, 4:7] = 0.5               # mapped direction (0,0,0) -> z = 0 -> theta = 0.5
    v[:, 3] = 0.625               # roughness at the centre of bin 2
    v[:, 7], v[:, 11], v[:, 12] = 0.1, 0.2, 0.3
    v[:, 13:16] = (0.4, 0.5, 0.6)
    x = M.encode(v, (-6, -4, -10), (6, 6, 2))
    pos_slots = [M.perm(0, g, j) for g in range(4) for j in range(8)] + [M.perm(1, g, 0) for g in range(4)]
    np.testing.assert_allclose(x[0, pos_slots], 0.0, atol=1e-12)
    # x = 0.5 (up to the FP32 1/extent): sin(2^k pi / 2) = 1 for k = 0, 0 for k >= 1
    assert abs(x[1, M.perm(0, 0, 0)] - 1.0) < 1e-6
    assert abs(x[1, M.perm(0, 1, 0)]) < 1e-6
    assert abs(x[0, M.perm(1, 2, 5)] - 1.0) < 1e-12           # roughness blob peak, bin 2
    assert abs(x[0, M.perm(1, 1, 5)] - np.exp(-0.5)) < 1e-12   # neighbouring bin
    np.testing.assert_allclose([x[0, M.perm(1, g, 6)] for g in range(3)], [0.1, 0.2, 0.3])
    np.testing.assert_allclose([x[0, M.perm(1, g, 7)] for g in range(3)], [0.4, 0.5, 0.6])
    assert x[0, M.perm(1, 3, 6)] == 1.0 and x[0, M.perm(1, 3, 7)] == 0.0
    # every logical input is produced exactly once
    assert sorted(M.perm(c, g, j) for c in range(2) for g in range(4) for j in range(8)) == list(range(64))


def test_adam_first_step_closed_form():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    w1, m, v, e = M.adam_step(w, np.zeros(3), np.zeros(3), w.copy(), g, 1, 0.01, 0.9, 0.99, 1e-8, 0.99)
    np.testing.assert_allclose(w1, w - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_allclose(e, 0.99 * w + 0.01 * w1, rtol=1e-12)


def test_targets_and_sanitize_special_values():
    rec = np.zeros((5, 11, 4))
    rec[:, 10, :3] = 0.5                                  # normaliser
    rec[0, 8, :3] = (1.0, 2.0, 3.0)                       # plain division
    # fminf(+inf, 1e5) = 1e5, fminf(NaN, 1e5) = 1e5 (C semantics), -inf -> 0
    rec[1, 8, :3] = (np.inf, -np.inf, np.nan)
    rec[2, 8, :3] = 1e9                                   # clamp to 1e5
    rec[3, 10, :3] = 0.0                                  # normaliser floor 1e-8
    rec[3, 8, :3] = 1e-9
    rec[4, 9, :3] = 2.0                                   # tail throughput x prediction
    pred = np.array([[0, 0, 0]] * 4 + [[1.0, -1.0, np.nan]], np.float64)
    t = M.targets(rec.reshape(5, 44), pred)
    np.testing.assert_allclose(t[0], [2.0, 4.0, 6.0])
    np.testing.assert_array_equal(t[1], [1e5, 0.0, 1e5])
    np.testing.assert_array_equal(t[2], [1e5, 1e5, 1e5])
    np.testing.assert_allclose(t[3], [0.1, 0.1, 0.1])
    np.testing.assert_allclose(t[4], [4.0, 0.0, 0.0])
    np.testing.assert_array_equal(M.sanitize([-1.0, np.nan, 2e5, np.inf, -np.inf, 3.0]), [0, 0, 1e5, 1e5, 0, 3.0])
