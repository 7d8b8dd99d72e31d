"""The learned-map MLP kernels: MFMA lane layout, gradients against the FP64
oracle, inference, and that Adam training reduces the loss."""
import numpy as np
import pytest

import oracle
from icempm_amd.learned import MLP, DeviceArray, init_params, mfma_probe, train_return_map

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [16, 32, 64])
def test_mfma_tile_layout(gpu_available, K):
    """The tile routine's lane layout: 16x16x16 steps for K = 16, gfx950's
    16x16x32 form for K = 32 and 64 (small integers: every product and sum is
    exact in BF16 operands with FP32 accumulation)."""
    print("mfma probe", K, flush=True)
    rng = np.random.default_rng(K)
    A = rng.integers(-8, 8, size=(16, K)).astype(np.float32)
    B = rng.integers(-8, 8, size=(K, 16)).astype(np.float32)
    D = mfma_probe(A, B)
    assert np.array_equal(D, A @ B)


def data(rng, n):
    X = rng.normal(size=(n, 5)).astype(np.float32)
    T = np.stack([np.tanh(X[:, 0] + 0.5 * X[:, 1]), X[:, 2] * X[:, 3] * 0.3,
                  np.sin(X[:, 4])], axis=1).astype(np.float32)
    return X, T


def test_gradient_against_fp64_oracle(gpu_available):
    print("mlp gradient", flush=True)
    rng = np.random.default_rng(2)
    P = init_params(3)
    X, T = data(rng, 1000)
    m = MLP(P)
    dX, dT = DeviceArray(X), DeviceArray(T)
    loss, g = m.step(dX, dT, 0, 1000, 0.0, want_grad=True)
    L, gref, _ = oracle.mlp_loss_grad(P.astype(np.float64), X, T)
    assert loss == pytest.approx(L, rel=2e-2)
    # BF16 operands: agreement to about 1e-2 of the gradient's norm
    assert np.linalg.norm(g - gref) / np.linalg.norm(gref) <= 3e-2
    for o, n in ((0, 1024), (1088, 4096), (5248, 1024)):
        blk, ref = g[o:o + n], gref[o:o + n]
        assert np.linalg.norm(blk - ref) / np.linalg.norm(ref) <= 5e-2
    # the parameters did not move (lr = 0)
    assert np.array_equal(m.params(), P)
    m.close()


def test_inference_against_fp64_oracle(gpu_available):
    print("mlp infer", flush=True)
    rng = np.random.default_rng(4)
    P = init_params(5)
    X, T = data(rng, 777)
    m = MLP(P)
    Y = m.infer(DeviceArray(X), 777)
    _, _, Yref = oracle.mlp_loss_grad(P.astype(np.float64), X, T, want_grad=False)
    assert np.abs(Y - Yref).max() <= 2e-2 * max(1.0, np.abs(Yref).max())
    m.close()


def test_training_reduces_loss(gpu_available):
    print("mlp train", flush=True)
    rng = np.random.default_rng(6)
    X, T = data(rng, 65536)
    m = MLP(init_params(7))
    dX, dT = DeviceArray(X), DeviceArray(T)
    first, _ = m.step(dX, dT, 0, 65536, 0.0)
    for it in range(400):
        r0 = (it * 4096) % 65536
        m.step(dX, dT, r0, 4096, 3e-3)
    last, _ = m.step(dX, dT, 0, 65536, 0.0)
    assert last < first / 10, (first, last)
    m.close()


def bf16(a):
    """float32 -> bfloat16 -> float32, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def test_inference_matches_bf16_emulation(gpu_available):
    """BF16 operands (inputs, weights, hidden activations), FP32 sums and
    biases: a numpy emulation agrees to FP32 summation-order differences."""
    print("mlp infer bf16", flush=True)
    rng = np.random.default_rng(8)
    P = init_params(9)
    P[1024:1088] = rng.normal(0, 0.1, 64)
    P[5184:5248] = rng.normal(0, 0.1, 64)
    P[6272:6275] = rng.normal(0, 0.1, 3)
    X, _ = data(rng, 333)
    m = MLP(P)
    Y = m.infer(DeviceArray(X), 333)
    m.close()
    f = np.float32
    W1 = bf16(P[0:1024].reshape(64, 16)[:, :5])
    W2 = bf16(P[1088:5184].reshape(64, 64))
    W3 = bf16(P[5248:6272].reshape(16, 64)[:3])
    h1 = bf16(np.tanh((bf16(X) @ W1.T).astype(f) + P[1024:1088].astype(f)))
    h2 = bf16(np.tanh((h1 @ W2.T).astype(f) + P[5184:5248].astype(f)))
    ref = (h2 @ W3.T).astype(f) + P[6272:6275].astype(f)
    # a hidden activation within an FP32 rounding of a BF16 tie can round the
    # other way; allow a few such flips
    err = np.abs(Y - ref)
    assert np.quantile(err, 0.99) <= 1e-4 * max(1.0, np.abs(ref).max())
    assert err.max() <= 2e-2 * max(1.0, np.abs(ref).max())


def test_training_driver(gpu_available):
    """icempm_mlp_train (the library's training driver, which
    learned.train_return_map calls): the normalisation it fits is the
    population mean and standard deviation of every column (a constant input
    column gets mean 0 and scale 1), training lowers the loss, and a seed
    fixes the result bit for bit (the gradient partials are summed in a fixed
    order)."""
    print("mlp training driver", flush=True)
    rng = np.random.default_rng(7)
    X, T = data(rng, 20000)
    X[:, 3] = 0.25                               # a constant feature
    p1, norm, losses = train_return_map(X, T, steps=400, batch=2048, lr=3e-3, seed=11, log_every=100)
    xm, xs = X.astype(np.float64).mean(0), X.astype(np.float64).std(0)
    xm[3], xs[3] = 0.0, 1.0
    tm, ts = T.astype(np.float64).mean(0), T.astype(np.float64).std(0)
    ref = np.concatenate([xm, xs, tm, ts])
    assert np.allclose(norm, ref, rtol=1e-5, atol=1e-6), (norm, ref)
    print("losses", losses, flush=True)
    assert [s for s, _ in losses] == [0, 100, 200, 300, 400]
    assert losses[-1][1] < 0.5 * losses[0][1]
    p2, norm2, _ = train_return_map(X, T, steps=400, batch=2048, lr=3e-3, seed=11)
    assert np.array_equal(p1.view(np.uint32), p2.view(np.uint32))
    assert np.array_equal(norm, norm2)
    p3, _, _ = train_return_map(X, T, steps=400, batch=2048, lr=3e-3, seed=12)
    assert not np.array_equal(p1, p3)
