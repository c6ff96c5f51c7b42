"""Surrogate MLP kernels against the FP64 oracle (whose gradients are pinned by finite
differences in test_oracle.py).

FP32 mode: forward within 1e-5 relative, weight gradients within 1e-4 (relative L2 per
layer).  BF16 mode (BF16 weights and activations, FP32 accumulation): forward within 2e-2
and gradients within 5e-2 of the oracle run on the same BF16-rounded weights.
"""
import numpy as np
import pytest

import oracle
from rscape_amd.mlp import MLP

pytestmark = pytest.mark.gpu
DIMS = (48, 128, 128, 128, 2)


def bf16_round(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def rand_params(rng, dims, scale=1.0):
    W = [rng.normal(0, scale / np.sqrt(dims[l]), (dims[l + 1], dims[l])).astype(np.float32) for l in range(len(dims) - 1)]
    b = [rng.normal(0, 0.1, dims[l + 1]).astype(np.float32) for l in range(len(dims) - 1)]
    return W, b


@pytest.mark.parametrize("B", [1, 31, 32, 100, 1000])
def test_forward_fp32(ctx, B):
    rng = np.random.default_rng(B)
    net = MLP(ctx, DIMS, bf16=False)
    W, b = rand_params(rng, DIMS)
    net.set_params(W, b)
    x = rng.normal(size=(B, DIMS[0])).astype(np.float32)
    y = net.forward(x)
    ref = oracle.mlp_forward_backward(W, b, x)
    assert relerr(y, ref) < 1e-5


@pytest.mark.parametrize("dims", [DIMS, (16, 32, 3), (32, 64, 64, 1), (48, 200, 2)])
def test_gradients_fp32(ctx, dims):
    rng = np.random.default_rng(len(dims))
    net = MLP(ctx, dims, bf16=False)
    W, b = rand_params(rng, dims)
    net.set_params(W, b)
    B = 77
    x = rng.normal(size=(B, dims[0])).astype(np.float32)
    t = rng.normal(size=(B, dims[-1])).astype(np.float32)
    dW, db, loss = net.grad(x, t)
    y = oracle.mlp_forward_backward(W, b, x)
    dy = 2.0 * (y - t) / (B * dims[-1])
    _, rW, rb = oracle.mlp_forward_backward(W, b, x, dy)
    assert np.isclose(loss, np.mean((y - t) ** 2), rtol=1e-5)
    for l in range(len(W)):
        assert relerr(dW[l], rW[l]) < 1e-4, l
        assert relerr(db[l], rb[l]) < 1e-4, l


def test_bf16_forward_and_gradients(ctx):
    rng = np.random.default_rng(5)
    net = MLP(ctx, DIMS, bf16=True)
    W, b = rand_params(rng, DIMS)
    net.set_params(W, b)
    B = 256
    x = rng.normal(size=(B, DIMS[0])).astype(np.float32)
    t = rng.normal(size=(B, DIMS[-1])).astype(np.float32)
    Wr = [bf16_round(w) for w in W]
    y = net.forward(x)
    yref = oracle.mlp_forward_backward(Wr, b, x)
    assert relerr(y, yref) < 2e-2
    dW, db, loss = net.grad(x, t)
    dy = 2.0 * (yref - t) / (B * DIMS[-1])
    _, rW, rb = oracle.mlp_forward_backward(Wr, b, x, dy)
    for l in range(len(W)):
        assert relerr(dW[l], rW[l]) < 5e-2, (l, relerr(dW[l], rW[l]))
        assert relerr(db[l], rb[l]) < 5e-2, l
    # parameters round-trip unchanged in FP32
    W2, b2 = net.get_params()
    assert all(np.array_equal(u, v) for u, v in zip(W, W2)) and all(np.array_equal(u, v) for u, v in zip(b, b2))


def target_fn(x):
    return np.stack([np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2], np.tanh(x[:, 3] - x[:, 4])], 1).astype(np.float32)


@pytest.mark.parametrize("bf16", [True, False])
def test_training_fits_and_is_deterministic(ctx, bf16):
    rng = np.random.default_rng(11)
    x = rng.normal(size=(4096, DIMS[0])).astype(np.float32)
    t = target_fn(x)
    net = MLP(ctx, DIMS, seed=3, bf16=bf16)
    l0 = float(np.mean((net.forward(x) - t) ** 2))
    loss = net.train(x, t, epochs=30, batch=256, lr=2e-3, seed=9)
    l1 = float(np.mean((net.forward(x) - t) ** 2))
    print("bf16", bf16, "initial", l0, "train", loss, "final", l1)
    assert l1 < 0.1 * l0 and l1 < 0.05
    net2 = MLP(ctx, DIMS, seed=3, bf16=bf16)
    net2.train(x, t, epochs=30, batch=256, lr=2e-3, seed=9)
    for u, v in zip(net.get_params()[0], net2.get_params()[0]):
        assert np.array_equal(u, v)
