"""The asynchronous closure schedule (C3D_LAUNCH_ASYNC) in a HIP graph.

c3d_update's SSP-RK2 step (update.cpp:185-275: per stage the ghost fill, the closure, the
kinetic flux and the Euler step, then the RK2 average) is captured once into a graph and
replayed; every replay must give, bit for bit, the state and closure status that the same
step gives when run eagerly.  Between replays the test runs eager work of its own (a torch
reduction) and synchronises.  Reference-order mode ties the eager steps to the oracle
(tests/test_gpu_grid.py); the mixed mode exercises the mixed sweep and its tail kernels.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("precision", ["reforder", "mixed"])
def test_update_step_replays_from_a_graph(P, O, dev, precision):
    import torch
    T2 = O.Tables(2, 8)
    n = 12
    d = 2.0 / n
    u, sT, sS = O.init_grid(n, n, n, T2.nm, ic=1, sigma=0.1, floor=1e-4, noise=0.01, seed=3,
                            cross=1.0)
    alpha = O.init_alpha(u)
    dt = 0.5 * (2 / 3) / (3 / d)
    prec = P.PREC_FP64_REFORDER if precision == "reforder" else P.PREC_MIXED
    opts = P.closure_opts()
    steps = 5

    def context():
        ctx = P.Ctx(tables=T2)
        ctx.set_precision(prec)
        ctx.set_tail(64, 8, wide_cells=16)
        ctx.set_launch(P.LAUNCH_ASYNC, 24)
        return ctx

    sTd, sSd = _dev(sT, dev), _dev(sS, dev)
    # eager steps
    ctx_e = context()
    ue, ae = _dev(u, dev), _dev(alpha, dev)
    eager = []
    for _ in range(steps):
        ctx_e.update(n, n, n, ue, ae, sTd, sSd, d, d, d, 1.0, dt, 1, 0, opts)
        rc = ctx_e.sync()
        eager.append((ue.cpu().numpy(), ae.cpu().numpy(), rc))
    ctx_e.close()
    # one eager step (sizes the context's buffers), then the captured step replayed
    ctx_g = context()
    ug, ag = _dev(u, dev), _dev(alpha, dev)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ctx_g.update(n, n, n, ug, ag, sTd, sSd, d, d, d, 1.0, dt, 1, 0, opts)
    assert ctx_g.sync(stream) == eager[0][2]
    assert np.array_equal(ug.cpu().numpy(), eager[0][0])
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        ctx_g.update(n, n, n, ug, ag, sTd, sSd, d, d, d, 1.0, dt, 1, 0, opts)
    for k in range(1, steps):
        graph.replay()
        torch.cuda.synchronize()
        checksum = float(ug.sum())  # eager work between replays
        assert np.isfinite(checksum)
        rc = ctx_g.sync()
        ue_k, ae_k, rc_k = eager[k]
        print(f"{precision} replay {k}: status {rc} (eager {rc_k})")
        assert rc == rc_k, k
        assert np.array_equal(ug.cpu().numpy(), ue_k), k
        assert np.array_equal(ag.cpu().numpy(), ae_k), k
    del graph
    ctx_g.close()
