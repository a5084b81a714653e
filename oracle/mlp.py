"""CPU oracle of the warm-start network (FP64 numpy) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this file,
and only as the checker.  The product (c3d-2015_amd/) never imports it.

The network has no counterpart in the reference: it is the learned warm start of the M_N
closure, a multilayer perceptron 25 -> 128 -> 128 -> 25 with exact (erf) GELU activations.
Parameter vector layout (23065 entries): W1[25][128] b1[128] W2[128][128] b2[128] W3[128][25]
b3[25]; a layer computes z_j = b_j + sum_i x_i W[i][j].  Loss: mean over the n rows and the
first nm outputs of (y - t)^2.  The gradient below is pinned by central finite differences
in tests/test_oracle_mlp.py.
"""
import math

import numpy as np

IN, HID = 25, 128
SHAPES = [("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, IN)), ("b3", (IN,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)

_erf = np.vectorize(math.erf)


def unpack(p):
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shape)
        o += k
    return out


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(p, X, cache=False):
    w = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    Z1 = X @ w["W1"] + w["b1"]
    A1 = gelu(Z1)
    Z2 = A1 @ w["W2"] + w["b2"]
    A2 = gelu(Z2)
    Y = A2 @ w["W3"] + w["b3"]
    if cache:
        return Y, (X, Z1, A1, Z2, A2)
    return Y


def backward(p, cache, dY):
    """Parameter gradient (flat) from the gradient dY with respect to the outputs."""
    w = unpack(p)
    X, Z1, A1, Z2, A2 = cache
    g = {"W3": A2.T @ dY, "b3": dY.sum(0)}
    dZ2 = (dY @ w["W3"].T) * gelu_d(Z2)
    g["W2"], g["b2"] = A1.T @ dZ2, dZ2.sum(0)
    dZ1 = (dZ2 @ w["W2"].T) * gelu_d(Z1)
    g["W1"], g["b1"] = X.T @ dZ1, dZ1.sum(0)
    return np.concatenate([g[name].ravel() for name, _ in SHAPES])


def loss_grad(p, X, T, nm):
    """Mean squared error over rows and outputs k < nm, and its gradient (flat)."""
    Y, cache = forward(p, X, cache=True)
    n = Y.shape[0]
    D = np.zeros_like(Y)
    D[:, :nm] = Y[:, :nm] - np.asarray(T, dtype=np.float64)[:, :nm]
    loss = float((D ** 2).sum() / (n * nm))
    return loss, backward(p, cache, 2.0 * D / (n * nm))


def dual_objective(alpha, ut, M, w):
    """u0-scaled dual objective f = sum_q w_q e^{min(alpha.m_q, 80)} - alpha.u~ and its
    gradient sum_q w_q e m_q - u~ (the moment residual), rows of alpha [n][nm], ut [n][nm]."""
    e = np.exp(np.minimum(alpha @ M.T, 80.0)) * w
    return e.sum(1) - (alpha * ut).sum(1), e @ M - ut


def dual_loss_grad(p, X, u, M, w, norm):
    """Mean over rows of f(alpha_mean + alpha_std * net(X)) for the u0-scaled moments of u,
    and its parameter gradient (flat)."""
    nm = M.shape[1]
    Y, cache = forward(p, X, cache=True)
    ym = np.asarray(norm["alpha_mean"], dtype=np.float64)[:nm]
    ys = np.asarray(norm["alpha_std"], dtype=np.float64)[:nm]
    alpha = ym + ys * Y[:, :nm]
    ut = np.asarray(u, dtype=np.float64) / np.asarray(u, dtype=np.float64)[:, :1]
    f, g = dual_objective(alpha, ut, M, w)
    n = Y.shape[0]
    dY = np.zeros_like(Y)
    dY[:, :nm] = g * ys / n
    return float(f.mean()), backward(p, cache, dY)


def features(u, nm, norm):
    """x_k = (u_k/u0 - x_mean_k)/x_std_k for k < nm, zero beyond (u: [n][nm])."""
    u = np.asarray(u, dtype=np.float64)
    X = np.zeros((u.shape[0], IN))
    X[:, :nm] = (u / u[:, :1] - norm["x_mean"][:nm]) / norm["x_std"][:nm]
    return X
