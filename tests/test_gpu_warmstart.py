"""What the learned warm start gains over the plain seeded guess: a small
coordinate MLP trained on the GPU (BF16 kernels) from converged slab
solutions predicts held-out instances; Newton from the prediction needs
fewer steps to the same absolute accuracy and reaches the plain FP64 fixed
point; a useless prediction falls back to the plain guess."""
import importlib.util
import os
import sys

import numpy as np
import pytest

from product_util import PKG, glhip

pytestmark = pytest.mark.gpu
G = glhip()
_spec = importlib.util.spec_from_file_location("warmstart", os.path.join(PKG, "warmstart.py"))
WS = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(WS)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def trained():
    train = WS.slab_instances(G, 1024, 3)
    Xs, Ys = [], []
    for b0 in range(0, len(train), 256):
        part = train[b0:b0 + 256]
        guesses = [WS.plain_guess(G, it, 3000 + b0 + i) for i, it in enumerate(part)]
        st, ist, F, Us = WS.solve_batch(G, part, guesses, fp32=True, rtol=1e-8)
        X, Y = WS.training_set(part, Us, ist[:, 1] == 1)
        Xs.append(X)
        Ys.append(Y)
        say("training batch", b0, int(ist[:, 1].sum()))
    net = G.MLP(n_freq=64, sigma=0.5, width=256, depth=5, n_out=18, seed=3)
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    net.set_data(X, Y)
    # the mean squared error over a fixed sample of the training data (the
    # sampled mini-batch losses of Adam fluctuate from step to step)
    sub = np.random.default_rng(3).choice(X.shape[0], min(32768, X.shape[0]), replace=False)
    mse = lambda: float(np.mean((net.predict(X[sub]) - Y[sub]) ** 2))
    e0 = mse()
    L = WS.train(net, 6000, 32768, lr=1e-3, segment=500, seed=3)
    e1 = mse()
    say("loss", L, "sample MSE", e0, "->", e1)
    assert e1 < 0.1 * e0
    yield net
    net.close()


def test_learned_guess_needs_fewer_newton_steps(trained):
    test = WS.slab_instances(G, 64, 4)
    plain = [WS.plain_guess(G, it, 4000 + i) for i, it in enumerate(test)]
    learned = WS.predict_guesses(trained, test)
    st_p, ist_p, F_p, res_p, _, rend_p = WS.solve_from_candidates(G, test, [plain], fp32=True)
    st_w, ist_w, F_w, res_w, choice, rend_w = WS.solve_from_candidates(G, test, [plain, learned], fp32=True)
    st_64, ist_64, F_64, _, _, rend_64 = WS.solve_from_candidates(G, test, [plain], fp32=False)
    say("plain", st_p, "mean steps", ist_p[:, 0].mean())
    say("learned", st_w, "mean steps", ist_w[:, 0].mean(), "used", int((choice == 1).sum()))
    assert np.all(ist_w[:, 1] == 1) and np.all(ist_p[:, 1] == 1) and np.all(ist_64[:, 1] == 1)
    # every reported convergence holds for the residual recomputed over the batch
    for rend in (rend_p, rend_w, rend_64):
        assert np.all(rend <= 1e-10 * res_p[0] * (1 + 1e-6))
    # the plain FP32 path reaches the FP64 fixed point from the same guess
    assert np.mean(np.abs(F_p - F_64) <= 1e-10 * np.abs(F_64)) >= 0.95
    assert (choice == 1).sum() >= 48                        # the prediction beats the plain guess mostly
    assert np.median(res_w[1]) < 0.5 * np.median(res_w[0])  # lower first residual
    assert ist_w[:, 0].mean() < ist_p[:, 0].mean() - 0.5    # fewer Newton steps per instance
    same = np.abs(F_w - F_64) <= 1e-10 * np.abs(F_64)
    assert same.mean() >= 0.95


def test_useless_prediction_falls_back_to_plain_guess(trained):
    test = WS.slab_instances(G, 16, 5)
    plain = [WS.plain_guess(G, it, 5000 + i) for i, it in enumerate(test)]
    junk = [10.0 * it.gap_B * np.random.default_rng(i).standard_normal(p.size) for i, (it, p) in
            enumerate(zip(test, plain))]
    st, ist, F, res, choice, rend = WS.solve_from_candidates(G, test, [plain, junk], fp32=True)
    assert np.all(choice == 0) and np.all(ist[:, 1] == 1)
