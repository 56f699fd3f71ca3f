"""The CPU reference of the coordinate MLP (oracle/mlp_ref.py), pinned before
it judges the GPU kernels: BF16 rounding against exact cases, the backward
pass against central finite differences of the loss in float64, and the
forward pass against an independent per-sample evaluation."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref as R  # noqa: E402


def test_bf16_rounding_cases():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, 1.0 + 2.0 ** -7, -2.5, 3.0e-5, 65504.0], np.float32)
    y = R.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-9 rounds up to 1 + 2^-7
    assert y[0] == 1.0 and y[1] == 1.0 and y[2] == np.float32(1.0 + 2.0 ** -7) and y[3] == y[2]
    assert y[4] == -2.5
    assert abs(y[5] - 3.0e-5) <= 2.0 ** -8 * 3.0e-5
    assert np.all(np.abs(y - x) <= np.abs(x) * 2.0 ** -8)


@pytest.mark.parametrize("depth", [1, 3])
def test_backward_matches_finite_differences(depth):
    F, W, n_out = 4, 8, 3
    K, N, offs, n_p = R.layout(F, W, depth, n_out)
    rng = np.random.default_rng(depth)
    p = rng.standard_normal(n_p) * 0.5
    B = rng.standard_normal((6, F))
    X = rng.uniform(-1, 1, (7, 6))
    Y = rng.standard_normal((7, n_out))
    L, g = R.loss_grad(p, B, X, Y, F, W, depth, n_out, emulate=False)
    for i in rng.choice(n_p, 40, replace=False):
        e = np.zeros(n_p)
        e[i] = 1e-6
        Lp, _ = R.loss_grad(p + e, B, X, Y, F, W, depth, n_out, emulate=False)
        Lm, _ = R.loss_grad(p - e, B, X, Y, F, W, depth, n_out, emulate=False)
        assert g[i] == pytest.approx((Lp - Lm) / 2e-6, rel=1e-6, abs=1e-9)
    # the loss is the mean squared error of the forward pass
    out = R.predict(p, B, X, F, W, depth, n_out, emulate=False)
    assert L == pytest.approx(np.mean((out - Y) ** 2), rel=1e-14)


def test_forward_matches_per_sample_evaluation():
    F, W, depth, n_out = 4, 8, 2, 3
    K, N, offs, n_p = R.layout(F, W, depth, n_out)
    rng = np.random.default_rng(5)
    p = rng.standard_normal(n_p)
    B = rng.standard_normal((6, F))
    X = rng.uniform(-1, 1, (5, 6))
    out = R.predict(p, B, X, F, W, depth, n_out, emulate=False)
    for s in range(5):
        feats = []
        for j in range(F):
            feats.append(np.sin(2 * np.pi * sum(X[s, d] * B[d, j] for d in range(6))))
        for j in range(F):
            feats.append(np.cos(2 * np.pi * sum(X[s, d] * B[d, j] for d in range(6))))
        h = feats
        for l in range(depth + 1):
            a, b = offs[l]
            Wl = p[a:b].reshape(K[l], N[l])
            bl = p[b:b + N[l]]
            z = [sum(h[i] * Wl[i, o] for i in range(K[l])) + bl[o] for o in range(N[l])]
            h = [zz / (1 + np.exp(-zz)) for zz in z] if l < depth else z
        assert np.allclose(out[s], h, rtol=1e-12, atol=1e-12)
