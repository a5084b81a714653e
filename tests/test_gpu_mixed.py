"""GPU checks of the mixed-precision closure (C3D_PREC_MIXED): the Hessian on the FP16
matrix cores as three split products under a log-sum-exp shift, the gradient summed in FP64
(FP64 matrix cores), exponents, f, Cholesky and condition indicator in FP64 and the FP64
stopping test in the mixed phase itself; ill-conditioned cells continue on FP64 Hessians
from their iterate near the optimum, or restart from their input on the FP64 path.

Bars, against the oracle (oracle/c3d_oracle.c, FP64 restatement of src/opt/*) and the
product's own FP64 mode on identical inputs:
  * the mixed Hessian is within 1e-5 of the FP64 one (relative to max|H|), also for
    exponents far from zero (its own shift keeps the split operands in FP16's range);
  * every converged cell meets the FP64 stopping test (||grad h||_2 <= tol in FP64,
    u0-scaled) and reproduces its moments to 1.5e-8 u0;
  * a cell handed to the FP64 path ends bit-identical to the FP64 mode's result;
  * the regularisation rung agrees with the oracle's, and the moments written back agree
    with the oracle's to 1e-6 u0, on >= 99.5% of the cells (moments of cells whose rung
    differs move by r >= 1e-3, optBatch.cpp:229-231).
"""
import numpy as np
import pytest

from test_gpu_closure import KEYS, TAILS, _run_gpu_closure, _set_tail, hess_lmin, moment_residual

pytestmark = pytest.mark.gpu


def _run(P, dev, T, u, a, opts, precision, tail=None, cond_max=None):
    import torch
    ctx = P.Ctx(tables=T)
    ctx.set_precision(precision)
    if cond_max is not None:
        ctx.set_mixed(cond_max)
    _set_tail(ctx, tail, u.shape[0])
    ud = torch.from_numpy(u.copy()).to(dev)
    ad = torch.from_numpy(a.copy()).to(dev)
    st = P.alloc_stats(u.shape[0], dev)
    ctx.closure(ud, ad, opts, stats=st)
    rc = ctx.sync()
    out = (ud.cpu().numpy(), ad.cpu().numpy(), P.stats_to_numpy(st), rc)
    ctx.close()
    return out


@pytest.mark.parametrize("N,Q", [(1, 6), (2, 8), (3, 24), (4, 12)])
def test_mixed_hessian_close_to_fp64(P, O, dev, N, Q):
    import torch
    T = O.Tables(N, Q)
    rng = np.random.default_rng(17 * N + Q)
    n = 41
    a = rng.normal(0, 0.7, (n, T.nm))
    a[::3, 0] += 60.0  # weights e^{+17}: scaled far beyond FP16's range without the shift
    a[1::3, 0] -= 60.0  # and e^{-17}: below its normal range
    h_ref = np.zeros((n, T.nm, T.nm))
    gx = np.zeros(T.nm2)
    for c in range(n):
        O.lib().orc_solveH(T.nm, T.nm2, T.nq, np.ascontiguousarray(a[c]), O.C.byref(T.c), 0, 0,
                           h_ref[c], gx)
    ctx = P.Ctx(tables=T)
    ctx.set_precision(P.PREC_MIXED)
    ad = torch.from_numpy(a).to(dev)
    hd = torch.zeros((n, T.nm, T.nm), dtype=torch.float64, device=dev)
    ctx.hessian(ad, hd)
    torch.cuda.synchronize()
    h = hd.cpu().numpy()
    ctx.close()
    assert np.array_equal(h, np.transpose(h, (0, 2, 1)))
    rel = np.abs(h - h_ref).max(axis=(1, 2)) / np.abs(h_ref).max(axis=(1, 2))
    print(f"max rel {rel.max():.2e}")
    assert rel.max() < 1e-5, rel.max()


@pytest.mark.parametrize("N,Q,n", [(3, 24, 3000), (2, 8, 2000), (1, 6, 1000), (4, 12, 600)])
def test_closure_mixed_against_oracle_and_fp64(P, O, dev, N, Q, n):
    T = O.Tables(N, Q)
    a_star, u, a0 = O.closure_batch_input(T, n, 0xC3D2015 + N)
    uo, ao = u.copy(), a0.copy()
    rc_o, so = O.closure(T, uo, ao, O.opts())
    um, am, sm, rc_m = _run(P, dev, T, u, a0, P.closure_opts(), P.PREC_MIXED)
    uf, af, sf, rc_f = _run(P, dev, T, u, a0, P.closure_opts(), P.PREC_FP64)
    fb = (sm["flags"] & P.FLAG_FP64_FALLBACK) != 0
    # the mixed phase solves nearly every cell itself (a fallback restarts from the input on
    # the FP64 path, whose results would pass every bar below)
    assert fb.mean() <= 0.01, fb.mean()
    # cells handed over end exactly as the FP64 mode ends them
    assert np.array_equal(am[fb], af[fb]) and np.array_equal(um[fb], uf[fb])
    for k in ("totalIter", "iterGamma", "regIndex", "r", "err", "gamma"):
        assert np.array_equal(sm[k][fb], sf[k][fb]), k
    assert np.array_equal(sm["flags"][fb] & ~P.FLAG_FP64_FALLBACK, sf["flags"][fb])
    # the other cells: FP64 stopping test, moments, tolerance ball
    mx = ~fb
    conv = mx & ((sm["flags"] & P.FLAG_LADDER) == 0)
    assert (sm["err"][conv] <= 1e-8).all()
    res = moment_residual(T, um[conv], am[conv])
    assert res.max() <= 1.5e-8, res.max()
    same_rung = sm["regIndex"] == so["regIndex"]
    ok = same_rung & (so["regIndex"] == 0) & conv
    dist = np.linalg.norm(am - ao, axis=1)
    ball = 2.5 * 1e-8 * np.abs(u[:, 0]) / np.maximum(hess_lmin(T, ao), 1e-300)
    assert (dist[ok] <= ball[ok]).all(), (dist / ball)[ok].max()
    # rungs and written-back moments against the oracle
    du = np.abs(um - uo).max(axis=1) / np.abs(u[:, 0])
    frac_rung = same_rung.mean()
    frac_mom = (du <= 1e-6).mean()
    it_m, it_f = sm["totalIter"][mx].mean(), sf["totalIter"].mean()
    fin = ((sm["flags"] & P.FLAG_FP64_FINISH) != 0).mean()
    print(f"fallback {fb.mean():.4f} finish {fin:.4f} rung-agree {frac_rung:.4f} "
          f"moments<=1e-6 {frac_mom:.4f} iters mixed {it_m:.2f} fp64 {it_f:.2f}")
    assert frac_rung >= 0.995 and frac_mom >= 0.995
    assert rc_m == rc_f


@pytest.mark.parametrize("N,Q,n", [(3, 24, 2000), (4, 12, 600)])
def test_closure_mixed_result_independent_of_worklist_strategy(P, O, dev, N, Q, n):
    T = O.Tables(N, Q)
    a_star, u, a0 = O.closure_batch_input(T, n, 0xC3D2015 + 77, 0.5, 8.0)
    outs = [_run(P, dev, T, u, a0, P.closure_opts(), P.PREC_MIXED, tail=t) for t in TAILS]
    assert ((outs[0][2]["flags"] & P.FLAG_FP64_FALLBACK) != 0).mean() <= 0.01
    for ug, ag, sg, rc in outs[1:]:
        assert rc == outs[0][3]
        assert np.array_equal(ag, outs[0][1]) and np.array_equal(ug, outs[0][0])
        for k in KEYS + ("r", "err", "gamma"):
            assert np.array_equal(sg[k], outs[0][2][k]), k


def test_closure_mixed_forced_handover(P, O, dev):
    """cond_max = 1 hands every cell to the FP64 path at its first iteration: the result is
    then the FP64 mode's, bit for bit, ladder and restarts included."""
    T = O.Tables(2, 8)
    a_star, u, a0 = O.closure_batch_input(T, 400, 5, 2.0, 6.0)
    um, am, sm, rc_m = _run(P, dev, T, u, a0, P.closure_opts(maxIter=3), P.PREC_MIXED,
                            cond_max=1.0)
    uf, af, sf, rc_f = _run(P, dev, T, u, a0, P.closure_opts(maxIter=3), P.PREC_FP64)
    assert (sm["flags"] & P.FLAG_FP64_FALLBACK).all()
    assert np.array_equal(am, af) and np.array_equal(um, uf) and rc_m == rc_f
    assert (sf["regIndex"] > 0).any()
