"""The FP64 restatement of the learned-map MLP: its analytic gradient
against central finite differences."""
import numpy as np

import oracle


def init_params(rng):
    p = np.zeros(oracle.MLP_PARAMS)
    p[0:1024] = rng.normal(0, 0.4, 1024)
    p[1088:5184] = rng.normal(0, 0.15, 4096)
    p[5248:6272] = rng.normal(0, 0.15, 1024)
    p[1024:1088] = rng.normal(0, 0.1, 64)
    p[5184:5248] = rng.normal(0, 0.1, 64)
    p[6272:6288] = rng.normal(0, 0.1, 16)
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    P = init_params(rng)
    X = rng.normal(size=(7, 5))
    T = rng.normal(size=(7, 3))
    L, g, _ = oracle.mlp_loss_grad(P, X, T)
    idx = rng.choice(oracle.MLP_PARAMS, 60, replace=False)
    # include parameters of every block
    idx = np.concatenate([idx, [3, 1030, 2000, 5190, 5250, 6273]])
    for i in idx:
        e = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += e
        Pm[i] -= e
        fd = (oracle.mlp_loss_grad(Pp, X, T, False)[0] - oracle.mlp_loss_grad(Pm, X, T, False)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])


def test_padding_gets_no_gradient():
    rng = np.random.default_rng(5)
    P = init_params(rng)
    _, g, _ = oracle.mlp_loss_grad(P, rng.normal(size=(4, 5)), rng.normal(size=(4, 3)))
    W1 = g[0:1024].reshape(64, 16)
    assert not W1[:, 5:].any()                      # inputs 5..15 are zero
    W3 = g[5248:6272].reshape(16, 64)
    assert not W3[3:].any() and not g[6275:6288].any()   # outputs 3..15 carry no loss


def test_forward_matches_numpy_layout():
    """The documented parameter layout (W[out][in], W1 64x16 at 0, b1 at 1024,
    W2 64x64 at 1088, b2 at 5184, W3 16x64 at 5248, b3 at 6272) evaluated
    with numpy matrix products."""
    rng = np.random.default_rng(6)
    P = init_params(rng)
    X = rng.normal(size=(9, 5))
    T = np.zeros((9, 3))
    _, _, Y = oracle.mlp_loss_grad(P, X, T, False)
    W1 = P[0:1024].reshape(64, 16)[:, :5]
    b1 = P[1024:1088]
    W2 = P[1088:5184].reshape(64, 64)
    b2 = P[5184:5248]
    W3 = P[5248:6272].reshape(16, 64)[:3]
    b3 = P[6272:6275]
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    ref = h2 @ W3.T + b3
    assert np.allclose(Y, ref, rtol=1e-12, atol=1e-12)
