"""The proposer network's HIP kernels against the FP64 oracle: outputs, loss
and gradients, deterministic training, and a learnable target."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _data(seed, N=300, nin=12, nout=5):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (N, nin)).astype(np.float32)
    T = rng.normal(0, 1, (N, nout)).astype(np.float32)
    M = (rng.random((N, nout)) < 0.8).astype(np.float32)
    return X, T, M


def test_gpu_forward_loss_and_gradient_match_the_oracle():
    from gmr.proposer import Mlp
    X, T, M = _data(1)
    net = Mlp(12, 64, 5, seed=3)
    p = net.get_params()
    loss_o, g_o, out_o = O.mlp_loss_grad(12, 64, 5, p.astype(np.float64), X, T, M)
    out = net.forward(X)
    np.testing.assert_allclose(out, out_o, rtol=1e-4, atol=1e-4)
    loss, g = net.loss_grad(X, T, M)
    assert abs(loss - loss_o) <= 1e-4 * abs(loss_o)
    rel = np.linalg.norm(g - g_o) / np.linalg.norm(g_o)
    print("gradient relative error", rel)
    assert rel <= 1e-4


def test_gpu_training_is_deterministic_and_learns():
    from gmr.proposer import Mlp
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (400, 12)).astype(np.float32)
    # a smooth target the network can fit
    T = np.stack([np.sin(X[:, 0]) + X[:, 1] * X[:, 2], X[:, 3] ** 2, np.abs(X[:, 4]), X[:, 5] - X[:, 6],
                  np.tanh(X[:, 7] * X[:, 8])], 1).astype(np.float32)
    M = np.ones_like(T)
    a, b = Mlp(12, 64, 5, seed=9), Mlp(12, 64, 5, seed=9)
    la = a.train(X, T, M, epochs=600, lr=3e-3)
    lb = b.train(X, T, M, epochs=600, lr=3e-3)
    assert np.array_equal(a.get_params(), b.get_params())
    assert la[-1] < 0.1 * la[0]
    # the trained parameters' loss agrees with the oracle's
    loss_o = O.mlp_loss_grad(12, 64, 5, a.get_params().astype(np.float64), X, T, M)[0]
    assert abs(a.loss_grad(X, T, M)[0] - loss_o) <= 1e-3 * loss_o + 1e-6
