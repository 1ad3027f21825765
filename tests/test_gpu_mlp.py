"""The cost model's HIP forward/backward against a float64 numpy
restatement of the same network (tolerances stated per check), and its
training on the GPU reducing the loss."""
import numpy as np
import pytest

import lcs_amd

pytestmark = pytest.mark.gpu


def unpack(p, I, H, O):
    o = 0
    W1 = p[o:o + H * I].reshape(H, I); o += H * I
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + O * H].reshape(O, H); o += O * H
    b3 = p[o:o + O]
    return W1, b1, W2, b2, W3, b3


def np_loss_grad(p, X, T, M, I, H, O):
    p = p.astype(np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p, I, H, O)
    z1 = X @ W1.T + b1; a1 = np.maximum(z1, 0)
    z2 = a1 @ W2.T + b2; a2 = np.maximum(z2, 0)
    y = a2 @ W3.T + b3
    cnt = M.sum()
    e = y - T
    loss = (M * e * e).sum() / cnt
    d3 = 2 * M * e / cnt
    gW3 = d3.T @ a2; gb3 = d3.sum(0)
    d2 = (d3 @ W3) * (z2 > 0)
    gW2 = d2.T @ a1; gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (z1 > 0)
    gW1 = d1.T @ X; gb1 = d1.sum(0)
    return y, loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def test_forward_and_gradient_match_numpy(ctx):
    I, H, O = 4, 64, 6
    mlp = lcs_amd.Mlp(ctx, I, H, O, seed=3)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(300, I)).astype(np.float32)
    T = rng.normal(size=(300, O)).astype(np.float32)
    M = (rng.random((300, O)) < 0.7).astype(np.float32)
    p = mlp.get_params()
    y_ref, loss_ref, g_ref = np_loss_grad(p, X.astype(np.float64), T, M, I, H, O)
    y = mlp.predict(X)
    # fp32 accumulation over <= 64 terms: relative 1e-4 of the output scale
    assert np.allclose(y, y_ref, rtol=1e-4, atol=1e-4 * np.abs(y_ref).max())
    loss, g = mlp.loss_grad(X, T, M)
    assert abs(loss - loss_ref) <= 1e-4 * abs(loss_ref)
    # gradients summed over 300 samples in fp32: 1e-3 of the largest entry
    assert np.allclose(g, g_ref, rtol=1e-3, atol=1e-3 * np.abs(g_ref).max())


def test_training_fits_a_smooth_cost_surface(ctx):
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, size=(512, 4)).astype(np.float32)
    T = np.stack([X[:, 0] * (k + 1) * 0.5 - X[:, 1] ** 2 + 0.1 * k for k in range(6)], 1).astype(np.float32)
    M = np.ones_like(T)
    mlp = lcs_amd.Mlp(ctx, 4, 64, 6, seed=2)
    _, g = mlp.loss_grad(X, T, M)
    loss0, _ = mlp.loss_grad(X, T, M)
    loss = mlp.train(X, T, M, steps=1500, lr=3e-3)
    assert loss < 0.02 * loss0 and loss < 0.01
    c = mlp.choose(X)
    assert np.array_equal(c, np.argmin(mlp.predict(X), 1))
