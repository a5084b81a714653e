"""GPU checks of the learned warm start (c3d-2015_amd/csrc/warmstart.hip and the closure's
selection kernel) against the CPU oracle of the network (oracle/mlp.py, FP64 numpy, pinned
by finite differences in test_oracle_mlp.py):
  * training kernels: loss and full gradient of the FP32 path within 1e-4 (relative to the
    largest entry) of the oracle's on the same FP32 inputs; the FP32 forward within 1e-5;
  * the FP16 matrix-core inference within 2e-2 of the oracle (FP16 weights/activations);
  * features within FP32 rounding;
  * training on solver-produced (u, alpha) pairs lowers the loss, and the closure started
    from the network needs fewer iterations than from the isotropic alpha while every cell
    still meets the FP64 stopping test.
"""
import numpy as np
import pytest

import mlp as MO
from test_gpu_closure import moment_residual

pytestmark = pytest.mark.gpu


def _t(x, dev, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(x))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(dev)


def _rand_params(rng, scale=0.1):
    return (rng.normal(0, scale, MO.NPARAM)).astype(np.float32)


def test_mlp_loss_grad_and_forward_match_oracle(P, O, dev):
    import torch
    T = O.Tables(2, 8)
    ctx = P.Ctx(tables=T)
    rng = np.random.default_rng(3)
    for n in (1, 63, 300, 1000):
        p = _rand_params(rng)
        X = rng.normal(size=(n, 25)).astype(np.float32)
        Tg = rng.normal(size=(n, 25)).astype(np.float32)
        pd, Xd, Td = _t(p, dev), _t(X, dev), _t(Tg, dev)
        gd = torch.zeros(MO.NPARAM, dtype=torch.float32, device=dev)
        ld = torch.zeros(1, dtype=torch.float32, device=dev)
        ctx.mlp_loss_grad(pd, Xd, Td, gd, ld)
        Yd = torch.zeros((n, 25), dtype=torch.float32, device=dev)
        ctx.mlp_forward(pd, Xd, Yd)
        torch.cuda.synchronize()
        loss, g = MO.loss_grad(p.astype(np.float64), X.astype(np.float64),
                               Tg.astype(np.float64), T.nm)
        Y = MO.forward(p.astype(np.float64), X.astype(np.float64))
        gg = gd.cpu().numpy().astype(np.float64)
        err_g = np.abs(gg - g).max() / np.abs(g).max()
        err_y = np.abs(Yd.cpu().numpy() - Y).max() / np.abs(Y).max()
        err_l = abs(float(ld.cpu()[0]) - loss) / loss
        print(f"n={n} grad {err_g:.2e} forward {err_y:.2e} loss {err_l:.2e}")
        assert err_g < 1e-4 and err_y < 1e-5 and err_l < 1e-5
    ctx.close()


def test_mlp_fp16_inference_and_features_match_oracle(P, O, dev):
    import torch
    T = O.Tables(3, 12)
    ctx = P.Ctx(tables=T)
    rng = np.random.default_rng(5)
    p = _rand_params(rng, 0.12)
    norm = {"x_mean": rng.normal(0, 0.1, 25).astype(np.float32),
            "x_std": rng.uniform(0.5, 2.0, 25).astype(np.float32),
            "alpha_mean": rng.normal(0, 1, 25).astype(np.float32),
            "alpha_std": rng.uniform(0.5, 3.0, 25).astype(np.float32)}
    ctx.mlp_set(p, norm)
    p_back, norm_back = ctx.mlp_get()
    assert np.array_equal(p_back, p)
    assert np.array_equal(norm_back.to_numpy()["alpha_std"], norm["alpha_std"])
    n = 1000 + 7
    a_star, u, a0 = O.closure_batch_input(T, n, 11, 0.5, 3.0)
    X = torch.zeros((n, 25), dtype=torch.float32, device=dev)
    ctx.mlp_features(_t(u, dev), norm, X)
    Xo = MO.features(u, T.nm, {k: v.astype(np.float64) for k, v in norm.items()})
    Xg = X.cpu().numpy()
    assert np.abs(Xg - Xo).max() <= 1e-5 * max(1.0, np.abs(Xo).max())
    Yd = torch.zeros((n, T.nm), dtype=torch.float64, device=dev)
    ctx.mlp_infer(X, Yd)
    torch.cuda.synchronize()
    y = MO.forward(p.astype(np.float64), Xg.astype(np.float64))[:, :T.nm]
    ref = norm["alpha_mean"][:T.nm] + norm["alpha_std"][:T.nm] * y
    err = np.abs(Yd.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"fp16 inference rel err {err:.2e}")
    assert err < 2e-2
    ctx.close()


def test_mlp_dual_loss_grad_matches_oracle(P, O, dev):
    import torch
    T = O.Tables(3, 12)
    ctx = P.Ctx(tables=T)
    rng = np.random.default_rng(9)
    n = 500
    a_star, u, a0 = O.closure_batch_input(T, n, 21, 0.5, 4.0)
    p = _rand_params(rng, 0.1)
    norm = {"x_mean": np.zeros(25, np.float32), "x_std": np.ones(25, np.float32),
            "alpha_mean": rng.normal(0, 0.3, 25).astype(np.float32),
            "alpha_std": rng.uniform(0.3, 1.0, 25).astype(np.float32)}
    X = MO.features(u, T.nm, norm).astype(np.float32)
    gd = torch.zeros(MO.NPARAM, dtype=torch.float32, device=dev)
    ld = torch.zeros(1, dtype=torch.float32, device=dev)
    ctx.mlp_dual_loss_grad(_t(p, dev), _t(X, dev), _t(u, dev), norm, gd, ld)
    torch.cuda.synchronize()
    loss, g = MO.dual_loss_grad(p.astype(np.float64), X.astype(np.float64), u, T.M, T.w,
                                {k: v.astype(np.float64) for k, v in norm.items()})
    gg = gd.cpu().numpy().astype(np.float64)
    err_g = np.abs(gg - g).max() / np.abs(g).max()
    err_l = abs(float(ld.cpu()[0]) - loss) / abs(loss)
    print(f"dual grad {err_g:.2e} loss {err_l:.2e}")
    assert err_g < 1e-4 and err_l < 1e-5
    ctx.close()


def _solved_pairs(P, ctx, n, seed, dev):
    import torch
    u = torch.empty((n, ctx.nm), dtype=torch.float64, device=dev)
    a = torch.empty_like(u)
    ctx.make_closure_batch(n, seed, 0.5, 8.0, u, a)
    ctx.closure(u, a, P.closure_opts())
    ctx.sync()
    return u, a


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_trained_warm_start_cuts_iterations(P, O, dev, precision):
    import torch
    T = O.Tables(3, 12)
    ctx = P.Ctx(tables=T)
    ut, at = _solved_pairs(P, ctx, 200_000, 0xA11CE, dev)
    l0, l1 = ctx.mlp_train(ut, at, mse_steps=1500, dual_steps=1500, batch=8192, lr=3e-3, seed=7)
    print(f"mse loss {l0:.4f} -> dual gap {l1:.4f}")
    assert l1 < 0.5 * l0
    # a fresh batch, solved from the isotropic alpha with and without the network
    n = 20_000
    u = torch.empty((n, ctx.nm), dtype=torch.float64, device=dev)
    a0 = torch.empty_like(u)
    ctx.make_closure_batch(n, 0xBEEF, 0.5, 8.0, u, a0)
    u_np, a0_np = u.cpu().numpy(), a0.cpu().numpy()
    if precision == "mixed":
        ctx.set_precision(P.PREC_MIXED)
        ctx.set_mixed(1e8)
    res = {}
    for warm in (False, True):
        ctx.set_warmstart(warm)
        ud, ad = u.clone(), a0.clone()
        st = P.alloc_stats(n, dev)
        ctx.closure(ud, ad, P.closure_opts(), stats=st)
        rc = ctx.sync()
        res[warm] = (ud.cpu().numpy(), ad.cpu().numpy(), P.stats_to_numpy(st), rc)
    ctx.close()
    s0, s1 = res[False][2], res[True][2]
    took = (s1["flags"] & P.FLAG_WARM) != 0
    it0, it1 = s0["totalIter"].mean(), s1["totalIter"].mean()
    fb = (s1["flags"] & P.FLAG_FP64_FALLBACK) != 0
    print(f"{precision}: proposal taken {took.mean():.3f} iters {it0:.2f} -> {it1:.2f} "
          f"fallback {fb.mean():.3f} rung-agree {(s0['regIndex'] == s1['regIndex']).mean():.4f}")
    assert took.mean() > 0.5 and it1 < 0.9 * it0
    conv = (s1["flags"] & P.FLAG_LADDER) == 0
    assert (s1["err"][conv] <= 1e-8).all()
    ug, ag = res[True][0], res[True][1]
    assert moment_residual(T, ug[conv], ag[conv]).max() <= 1.5e-8
    assert (s0["regIndex"] == s1["regIndex"]).mean() >= 0.99
    du = np.abs(ug - res[False][0]).max(axis=1) / np.abs(u_np[:, 0])
    assert (du <= 1e-6).mean() >= 0.99
