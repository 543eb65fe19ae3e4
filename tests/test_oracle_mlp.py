"""Pins the CPU surrogate-network oracle (oracle/mlp_ref.py): its analytic
gradient against central finite differences of its own loss, and its
bfloat16 rounding against hand-worked bit patterns."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import mlp_ref  # noqa: E402


def _layout(nin, H):
    sz = [H * nin, H, H * H, H, H * H, H, H, 1]
    off, o = [], 0
    for s in sz:
        off.append(o)
        o += (s + 7) & ~7
    return off, o


def test_bf16_rounding():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1) wins;
    # 1 + 3*2^-8 ties between 1 + 2^-7 (odd) and 1 + 2^-6: rounds up
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8,
                  1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0])
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 0.0])
    np.testing.assert_array_equal(mlp_ref.bf16(v), want)
    assert np.isnan(mlp_ref.bf16(np.array([np.nan])))[0]


def test_gelu_derivative():
    z = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (mlp_ref.gelu(z + h) - mlp_ref.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(mlp_ref.gelu_prime(z), fd, atol=1e-8)
    # GELU(0) = 0, GELU(z) -> z for large z, -> 0 for very negative z
    assert mlp_ref.gelu(np.array([0.0]))[0] == 0.0
    assert abs(mlp_ref.gelu(np.array([8.0]))[0] - 8.0) < 1e-12
    assert abs(mlp_ref.gelu(np.array([-8.0]))[0]) < 1e-12


def test_gradient_matches_finite_differences():
    nin, H, M = 3, 6, 5
    off, n = _layout(nin, H)
    rng = np.random.default_rng(4)
    w = rng.normal(scale=0.7, size=n)
    xs = rng.normal(size=(M, nin))
    ys = rng.normal(size=M)
    loss, g = mlp_ref.loss_grad(w, off, nin, H, xs, ys, emulate=False)
    h = 1e-6
    used = np.zeros(n, dtype=bool)
    for k, sz in enumerate([H * nin, H, H * H, H, H * H, H, H, 1]):
        used[off[k]:off[k] + sz] = True
    for i in np.flatnonzero(used):
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        lp, _ = mlp_ref.loss_grad(wp, off, nin, H, xs, ys, emulate=False)
        lm, _ = mlp_ref.loss_grad(wm, off, nin, H, xs, ys, emulate=False)
        fd = (lp - lm) / (2 * h)
        assert abs(g[i] - fd) <= 1e-7 + 1e-6 * abs(fd), (i, g[i], fd)
    # padding between blocks carries no gradient
    assert np.all(g[~used] == 0.0)


def test_forward_is_the_composed_network():
    """A one-unit hand evaluation: nin=1, H=1."""
    off, n = _layout(1, 1)
    w = np.zeros(n)
    vals = [0.5, 0.1, -1.5, 0.2, 2.0, -0.3, 1.25, 0.05]
    for k, v in enumerate(vals):
        w[off[k]] = v
    x = np.array([[0.8]])
    a = x[0, 0]
    for l in range(3):
        a = mlp_ref.gelu(np.array([vals[2 * l] * a + vals[2 * l + 1]]))[0]
    want = vals[6] * a + vals[7]
    out, _ = mlp_ref.forward(w, off, 1, 1, x, emulate=False)
    assert abs(out[0] - want) < 1e-15


def test_whiten_diagonal_is_standardize():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(50, 4)).astype(np.float32)
    m = rng.normal(size=4).astype(np.float32)
    sd = (0.3 + rng.random(4)).astype(np.float32)
    w = np.diag(np.float32(1.0) / sd)
    # equal to float32 rounding (whiten keeps the product in float64)
    np.testing.assert_allclose(mlp_ref.whiten(x, m, w),
                               mlp_ref.standardize(x, m, sd), rtol=1e-7)
    # a full matrix: the rows of W (x - m)
    W = rng.normal(size=(4, 4))
    z = mlp_ref.whiten(x, m, W)
    d = x.astype(np.float64) - m
    np.testing.assert_allclose(z[7], W.astype(np.float32) @ d[7], rtol=1e-6)
