"""float64 numpy restatement of the learned scheduler's network and its
cross-entropy gradient -- TEST INFRASTRUCTURE ONLY (the checker of the HIP
forward / backward kernels).

Network 12 -> 64 -> 64 -> 16, ReLU, softmax; parameter layout (6032 floats,
row-major (out, in)): W1[64][12] b1[64] W2[64][64] b2[64] W3[16][64] b3[16].
Loss: mean over samples of -log softmax(z)[y].

Pinned by central finite differences of `loss` (test_mlp_oracle.py): a wrong
transpose, a missing ReLU mask or a dropped bias term fails them.
"""
import numpy as np

NI, NH, NO = 12, 64, 16
SHAPES = [("W1", (NH, NI)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NO, NH)), ("b3", (NO,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(P):
    P = np.asarray(P, np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = P[o:o + n].reshape(shp)
        o += n
    return out


def forward(P, X):
    p = unpack(P)
    X = np.asarray(X, np.float64)
    a1 = X @ p["W1"].T + p["b1"]
    h1 = np.maximum(a1, 0)
    a2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(a2, 0)
    z = h2 @ p["W3"].T + p["b3"]
    zs = z - z.max(axis=1, keepdims=True)
    e = np.exp(zs)
    return e / e.sum(axis=1, keepdims=True), (X, a1, h1, a2, h2, z)


def loss(P, X, Y):
    probs, _ = forward(P, X)
    return float(-np.mean(np.log(probs[np.arange(len(Y)), Y])))


def loss_grad(P, X, Y):
    p = unpack(P)
    probs, (X, a1, h1, a2, h2, z) = forward(P, X)
    n = len(Y)
    dz = probs.copy()
    dz[np.arange(n), Y] -= 1.0
    dz /= n
    gW3 = dz.T @ h2
    gb3 = dz.sum(0)
    dh2 = (dz @ p["W3"]) * (a2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ p["W2"]) * (a1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return float(-np.mean(np.log(probs[np.arange(n), Y]))), g
