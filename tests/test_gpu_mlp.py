"""The coordinate-MLP kernels (BF16 matrix cores, FP32 accumulation) against
the CPU reference oracle/mlp_ref.py, which reproduces the BF16 rounding points
and accumulates in float64: forward pass, loss and gradient at small and at
the full training shapes, and a training run that lowers the loss."""
import os
import sys

import numpy as np
import pytest

from product_util import ROOT, glhip

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_ref as R  # noqa: E402

pytestmark = pytest.mark.gpu
G = glhip()


def say(*a):
    print(*a, flush=True)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def _layers(net, g):
    K, N, offs, n_p = R.layout(net.n_freq, net.width, net.depth, net.n_out)
    out = []
    for (a, b), n in zip(offs, N):
        out.append((g[a:b], g[b:b + n]))
    return out


@pytest.mark.parametrize("shape", [(32, 64, 2), (64, 256, 5)])
def test_forward_and_gradient_match_reference(shape):
    F, W, D = shape
    say("mlp shape", shape)
    net = G.MLP(n_freq=F, sigma=2.0, width=W, depth=D, n_out=18, seed=3)
    try:
        p, B = net.params()
        rng = np.random.default_rng(1)
        n = 200 if W == 64 else 1024
        X = rng.uniform(-1, 1, (n, 6)).astype(np.float32)
        Y = rng.standard_normal((n, 18)).astype(np.float32)
        out = net.predict(X)
        ref = R.predict(p, B, X, F, W, D, 18)
        e = _rel(out, ref)
        say("forward rel err", e)
        assert e <= 5e-3
        L, g = net.loss_grad(X, Y)
        Lr, gr = R.loss_grad(p, B, X, Y, F, W, D, 18)
        say("loss", L, Lr)
        assert L == pytest.approx(Lr, rel=2e-3)
        for l, ((gw, gb), (rw, rb)) in enumerate(zip(_layers(net, g), _layers(net, gr))):
            ew, eb = _rel(gw, rw), _rel(gb, rb)
            say("layer", l, "dW rel err", ew, "db rel err", eb)
            assert ew <= 2e-2 and eb <= 2e-2
    finally:
        net.close()


def test_training_lowers_the_loss():
    net = G.MLP(n_freq=32, sigma=0.2, width=128, depth=3, n_out=18, seed=0)
    try:
        rng = np.random.default_rng(2)
        X = rng.uniform(-1, 1, (20000, 6)).astype(np.float32)
        o = np.arange(18)
        Y = (np.sin(np.pi * X[:, :1] + o / 3.0) * np.cos(0.5 * np.pi * X[:, 2:3]) + 0.1 * X[:, 3:4]).astype(
            np.float32)
        net.set_data(X, Y)
        L = net.train(400, batch=4096, lr=1e-3, seed=5, loss_every=50)
        say("loss curve", L)
        assert L[-1] < 0.1 * L[0]
        Xt = rng.uniform(-1, 1, (2000, 6)).astype(np.float32)
        Yt = np.sin(np.pi * Xt[:, :1] + o / 3.0) * np.cos(0.5 * np.pi * Xt[:, 2:3]) + 0.1 * Xt[:, 3:4]
        err = np.sqrt(np.mean((net.predict(Xt) - Yt) ** 2))
        say("held-out rms", err)
        assert err < 0.15
    finally:
        net.close()
