"""The proposer network's FP64 oracle, pinned by central finite differences."""
import numpy as np

from oracle import oracle as O


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    nin, nh, nout, N = 5, 7, 3, 9
    npar = nh * nin + nh + nh * nh + nh + nout * nh + nout
    p = rng.normal(0, 0.5, npar)
    X = rng.normal(0, 1, (N, nin))
    T = rng.normal(0, 1, (N, nout))
    M = (rng.random((N, nout)) < 0.7).astype(float)
    loss, g, out = O.mlp_loss_grad(nin, nh, nout, p, X, T, M)
    assert loss > 0 and out.shape == (N, nout)
    eps = 1e-6
    fd = np.zeros_like(p)
    for i in range(npar):
        q = p.copy()
        q[i] += eps
        lp = O.mlp_loss_grad(nin, nh, nout, q, X, T, M)[0]
        q[i] -= 2 * eps
        lm = O.mlp_loss_grad(nin, nh, nout, q, X, T, M)[0]
        fd[i] = (lp - lm) / (2 * eps)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_mlp_oracle_closed_form_single_layer_case():
    """With W1 = I-like and all-positive activations the network is affine:
    check the output against the composed affine map."""
    nin = nh = 3
    nout = 2
    W1 = np.eye(3)
    b1 = np.full(3, 10.0)          # keeps every unit active
    W2 = np.eye(3) * 2.0
    b2 = np.zeros(3)
    W3 = np.array([[1.0, 0, 0], [0, 1.0, 1.0]])
    b3 = np.array([0.5, -1.0])
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])
    X = np.array([[1.0, 2.0, 3.0], [0.0, -1.0, 1.0]])
    T = np.zeros((2, 2))
    M = np.ones((2, 2))
    _, _, out = O.mlp_loss_grad(nin, nh, nout, p, X, T, M)
    expect = ((X + 10.0) * 2.0) @ W3.T + b3
    np.testing.assert_allclose(out, expect, rtol=1e-12)
