"""Pins the CPU oracle of the warm-start network (oracle/mlp.py): its gradient against
central finite differences of its own loss, GELU against the closed form, and the
parameter layout."""
import math

import numpy as np

import mlp as M


def test_parameter_count():
    assert M.NPARAM == 23065 == 25 * 128 + 128 + 128 * 128 + 128 + 128 * 25 + 25


def test_gelu_closed_form_and_derivative():
    z = np.linspace(-6, 6, 241)
    ref = np.array([0.5 * x * (1 + math.erf(x / math.sqrt(2))) for x in z])
    assert np.allclose(M.gelu(z), ref, rtol=0, atol=1e-15)
    assert M.gelu(np.array([0.0]))[0] == 0.0
    h = 1e-5
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.abs(fd - M.gelu_d(z)).max() < 1e-9


def test_forward_structure():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.1, M.NPARAM)
    w = M.unpack(p)
    X = rng.normal(size=(5, 25))
    # W3 = 0: the output is b3 whatever the input
    p2 = p.copy()
    o = 25 * 128 + 128 + 128 * 128 + 128
    p2[o:o + 128 * 25] = 0.0
    assert np.array_equal(M.forward(p2, X), np.tile(w["b3"], (5, 1)))
    # one row by hand
    x = X[2]
    a1 = [0.5 * z * (1 + math.erf(z / math.sqrt(2))) for z in x @ w["W1"] + w["b1"]]
    a2 = [0.5 * z * (1 + math.erf(z / math.sqrt(2))) for z in np.array(a1) @ w["W2"] + w["b2"]]
    y = np.array(a2) @ w["W3"] + w["b3"]
    assert np.allclose(M.forward(p, X)[2], y, rtol=1e-13, atol=1e-13)


def test_gradient_against_finite_differences():
    rng = np.random.default_rng(1)
    p = rng.normal(0, 0.2, M.NPARAM)
    X = rng.normal(size=(7, 25))
    T = rng.normal(size=(7, 25))
    nm = 9
    loss, g = M.loss_grad(p, X, T, nm)
    Y = M.forward(p, X)
    assert abs(loss - ((Y[:, :nm] - T[:, :nm]) ** 2).mean()) < 1e-15
    idx = np.concatenate([rng.choice(M.NPARAM, 60, replace=False),
                          [0, 3199, 3200, 3327, 3328, 19711, 19712, 19839, 19840, 23039, 23040,
                           23064]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.loss_grad(pp, X, T, nm)[0] - M.loss_grad(pm, X, T, nm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])
    # outputs k >= nm carry no gradient into W3's columns
    w3 = g[25 * 128 + 128 + 128 * 128 + 128:][:128 * 25].reshape(128, 25)
    assert np.all(w3[:, nm:] == 0.0)


def test_dual_loss_gradient_against_finite_differences(O):
    T = O.Tables(1, 4)
    rng = np.random.default_rng(2)
    p = rng.normal(0, 0.2, M.NPARAM)
    a_star, u, a0 = O.closure_batch_input(T, 6, 3, 0.5, 2.0)
    norm = {"x_mean": np.zeros(25), "x_std": np.ones(25),
            "alpha_mean": rng.normal(0, 0.3, 25), "alpha_std": rng.uniform(0.5, 1.5, 25)}
    X = M.features(u, T.nm, norm)
    loss, g = M.dual_loss_grad(p, X, u, T.M, T.w, norm)
    # the gradient of f with respect to alpha is the moment residual
    al = rng.normal(0, 0.5, (3, T.nm))
    ut = u[:3] / u[:3, :1]
    f0, g0 = M.dual_objective(al, ut, T.M, T.w)
    for k in range(T.nm):
        d = np.zeros_like(al)
        d[:, k] = 1e-6
        fd = (M.dual_objective(al + d, ut, T.M, T.w)[0] - M.dual_objective(al - d, ut, T.M, T.w)[0]) / 2e-6
        assert np.abs(fd - g0[:, k]).max() < 1e-7
    idx = np.concatenate([rng.choice(M.NPARAM, 40, replace=False), [0, 3200, 23040, 23043]])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (M.dual_loss_grad(pp, X, u, T.M, T.w, norm)[0] -
              M.dual_loss_grad(pm, X, u, T.M, T.w, norm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])
