"""c3d_closure_noalloc (the closure with all scratch in a fixed device array) against
c3d_closure on the same inputs, through the C-ABI: bit-identical u, alpha and per-cell
outcomes, on two calls in a row, the second with a smaller batch (the array keeps the first
call's contents), in the FP64, reference-order and mixed-with-warm-start modes; batches
beyond the array's capacity are rejected with C3D_ERR_SCRATCH and leave their inputs
untouched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _closure(P, ctx, fn, u, a, dev, opts=None):
    import torch
    ud = torch.from_numpy(u.copy()).to(dev)
    ad = torch.from_numpy(a.copy()).to(dev)
    st = P.alloc_stats(u.shape[0], dev)
    rc = fn(ctx.h, u.shape[0], P._p(ud), P._p(ad), P.C.byref(opts or P.closure_opts()),
            P._p(st), P._stream(None))
    sync = ctx.sync()
    torch.cuda.synchronize()
    return rc, sync, ud.cpu().numpy(), ad.cpu().numpy(), P.stats_to_numpy(st)


def _same(x, y):
    assert x[0] == y[0] == 0 and x[1] == y[1]
    assert np.array_equal(x[2], y[2]) and np.array_equal(x[3], y[3])
    for k in ("r", "err", "gamma", "totalIter", "iterGamma", "regIndex", "flags"):
        assert np.array_equal(x[4][k], y[4][k]), k


@pytest.mark.parametrize("N,Q,mode", [(3, 24, "fp64"), (2, 8, "reforder"), (3, 12, "mixed+nn"),
                                      (4, 12, "fp64")])
def test_noalloc_closure_matches_closure(P, O, dev, N, Q, mode):
    import torch
    L = P.load()
    T = O.Tables(N, Q)
    ctx = P.Ctx(tables=T)
    if mode == "reforder":
        ctx.set_precision(P.PREC_FP64_REFORDER)
    if mode == "mixed+nn":
        n = 20_000
        ut = torch.empty((n, T.nm), dtype=torch.float64, device=dev)
        at = torch.empty_like(ut)
        ctx.make_closure_batch(n, 0xA11CE, 0.5, 8.0, ut, at)
        ctx.closure(ut, at, P.closure_opts())
        ctx.sync()
        ctx.mlp_train(ut, at, mse_steps=200, dual_steps=200, batch=4096, lr=3e-3, seed=3)
        ctx.set_warmstart(True)
        ctx.set_precision(P.PREC_MIXED)
    for call, n in enumerate((3000, 700)):  # second call smaller: stale contents in the array
        _, u, a0 = O.closure_batch_input(T, n, 0xC3D2015 + 11 * call + N, 0.5, 8.0)
        ref = _closure(P, ctx, L.c3d_closure, u, a0, dev)
        got = _closure(P, ctx, L.c3d_closure_noalloc, u, a0, dev)
        _same(got, ref)
        print(f"N={N} Q={Q} {mode} call {call} n={n}: identical; mean iters "
              f"{ref[4]['totalIter'].mean():.2f}")
    ctx.close()


def test_noalloc_rejects_batches_beyond_the_scratch(P, O, dev):
    import torch
    L = P.load()
    T = O.Tables(4, 8)
    ctx = P.Ctx(tables=T)
    nm = T.nm
    per_cell = (8 * nm + 12 + nm * nm) * 8 + 44
    n = P.NOALLOC_SCRATCH_BYTES // per_cell + 1  # one cell too many
    u = torch.ones((n, nm), dtype=torch.float64, device=dev)
    a = torch.zeros_like(u)
    rc = L.c3d_closure_noalloc(ctx.h, n, P._p(u), P._p(a), P.C.byref(P.closure_opts()), None,
                               P._stream(None))
    assert rc == P.C3D_ERR_SCRATCH
    torch.cuda.synchronize()
    assert bool((u == 1).all()) and bool((a == 0).all())
    # a batch that fits runs, as c3d_closure runs it
    _, uu, aa = O.closure_batch_input(T, 64, 5, 0.5, 2.0)
    got = _closure(P, ctx, L.c3d_closure_noalloc, uu, aa, dev)
    ref = _closure(P, ctx, L.c3d_closure, uu, aa, dev)
    _same(got, ref)
    ctx.close()
