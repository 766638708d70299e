"""Pins the numpy scheduler-network oracle (CPU): analytic gradient against
central finite differences, and the softmax against a closed form."""
import numpy as np

from oracle import mlp_ref


def test_param_count():
    assert mlp_ref.NPARAMS == 6032


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = rng.normal(0, 0.3, mlp_ref.NPARAMS)
    X = rng.uniform(0, 1, (9, 12))
    Y = rng.integers(0, 16, 9)
    _, g = mlp_ref.loss_grad(P, X, Y)
    idx = np.concatenate([rng.choice(768, 15, replace=False), 768 + np.arange(0, 64, 9),
                          832 + rng.choice(4096, 15, replace=False), 4928 + np.arange(0, 64, 9),
                          4992 + rng.choice(1024, 15, replace=False), 6016 + np.arange(16)])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (mlp_ref.loss(Pp, X, Y) - mlp_ref.loss(Pm, X, Y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_softmax_closed_form():
    # all weights zero: z = b3, so the probabilities are softmax(b3) for any input
    P = np.zeros(mlp_ref.NPARAMS)
    P[6016:] = np.arange(16) * 0.1
    probs, _ = mlp_ref.forward(P, np.ones((2, 12)))
    e = np.exp(np.arange(16) * 0.1)
    assert np.allclose(probs, e / e.sum(), atol=1e-15)
