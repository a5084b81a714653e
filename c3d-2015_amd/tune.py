"""Launch-shape tuner of the closure (c3d_closure) for one GPU.

The closure's results do not depend on its launch shapes (tests/test_gpu_closure.py and
test_gpu_mixed.py check bulk, fused and switching worklist strategies bit for bit), only its
time does.  For every (moment order N, quadrature order Q, precision) asked for, this times a
seeded batch of the closure workload (alpha*_k ~ N(0, s^2), s ~ U[0.5, 8], Newton from the
isotropic alpha; with the mixed precision, from a briefly trained warm start, as the closure
runs in the benchmark) under each candidate of a small grid of:

  tail_cells   cells left when the passes switch to the fused tail kernel (c3d_set_tail)
  tail_passes  rounds per tail launch (c3d_set_tail)
  tail_wide    cells left when the tail evaluates four line-search trials per round
               (c3d_set_tail_wide)

and writes the fastest to a JSON table keyed "N,Q,precision", which Ctx.apply_tuning() (and
through it bench.py) reads.  Usage (on the GPU):

    python c3d-2015_amd/tune.py [--out c3d-2015_amd/tuning.json] [--cells 262144]
                                [--shapes 3,24 1,24 2,24 4,24 2,8 4,40] [--precisions fp64 mixed]

The tail thresholds are cell counts, so an entry fits batches of about its "cells" size; the
closure benchmark's shape (3,24) is tuned at its own batch size (--shapes 3,24 --cells
1048576, which updates that entry only).
"""
import argparse
import importlib.util
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_TABLE = os.path.join(HERE, "tuning.json")


def front_end():
    if "c3d_hip" in sys.modules:
        return sys.modules["c3d_hip"]
    spec = importlib.util.spec_from_file_location("c3d_hip", os.path.join(HERE, "c3d_hip.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["c3d_hip"] = mod
    spec.loader.exec_module(mod)
    return mod


def candidates(precision, nm):
    grid = {"tail_cells": (1024, 2048, 4096, 8192, 32768), "tail_passes": (32, 128),
            "tail_wide": (256, 512, 2048, 8192)}
    keys = list(grid)
    for vals in itertools.product(*(grid[k] for k in keys)):
        c = dict(zip(keys, vals))
        if c["tail_wide"] <= c["tail_cells"]:
            yield c


def tune_shape(P, torch, dev, N, Q, precision, cells, seed, reps, log):
    ctx = P.Ctx(N, Q)
    nm = ctx.nm
    u0 = torch.empty((cells, nm), dtype=torch.float64, device=dev)
    a0 = torch.empty_like(u0)
    ctx.make_closure_batch(cells, seed, 0.5, 8.0, u0, a0)
    opts = P.closure_opts()
    if precision == "mixed":
        # a briefly trained warm start, so that the iteration profile is the benchmark's
        m = min(4 * cells, 1 << 20)
        ut = torch.empty((m, nm), dtype=torch.float64, device=dev)
        at = torch.empty_like(ut)
        ctx.make_closure_batch(m, seed + 1, 0.5, 8.0, ut, at)
        ctx.closure(ut, at, opts)
        ctx.sync()
        ctx.mlp_train(ut, at, mse_steps=1000, dual_steps=4000, batch=8192, lr=2e-3, seed=seed)
        del ut, at
        ctx.set_precision(P.PREC_MIXED)
        ctx.set_warmstart(True)
    u, a = torch.empty_like(u0), torch.empty_like(a0)

    def run(c):
        ctx.set_tail(c["tail_cells"], c["tail_passes"], wide_cells=c["tail_wide"])
        best = float("inf")
        for r in range(reps + 1):
            u.copy_(u0)
            a.copy_(a0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.closure(u, a, opts)
            rc = ctx.sync()
            dt = time.perf_counter() - t0
            if r > 0:  # the first run warms the kernels up
                best = min(best, dt)
        return best, rc

    results = []
    for c in candidates(precision, nm):
        t, rc = run(c)
        results.append((t, c))
        log(f"  N={N} Q={Q} {precision} {c}: {1e3 * t:.2f} ms (status {rc})")
    results.sort(key=lambda x: x[0])
    best_t, best = results[0]
    default = next((t for t, c in results if c["tail_cells"] == 8192 and c["tail_passes"] == 32
                    and c["tail_wide"] == 512), None)
    ctx.close()
    return {**best, "ms": round(1e3 * best_t, 3),
            "default_ms": None if default is None else round(1e3 * default, 3),
            "cells": cells, "seed": hex(seed)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--cells", type=int, default=262144)
    ap.add_argument("--seed", type=lambda s: int(s, 0), default=0x7E57)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--shapes", nargs="*", default=["3,24", "1,24", "2,24", "4,24", "2,8", "4,40"])
    ap.add_argument("--precisions", nargs="*", default=["fp64", "mixed"],
                    choices=["fp64", "mixed"])
    args = ap.parse_args()
    import torch
    P = front_end()
    dev = torch.device("cuda:0")
    log = lambda m: print(m, file=sys.stderr, flush=True)
    table = {}
    if os.path.exists(args.out):
        table = json.load(open(args.out))
    for shape in args.shapes:
        N, Q = (int(x) for x in shape.split(","))
        for prec in args.precisions:
            r = tune_shape(P, torch, dev, N, Q, prec, args.cells, args.seed, args.reps, log)
            table[f"{N},{Q},{prec}"] = r
            log(f"N={N} Q={Q} {prec}: best {r}")
            with open(args.out, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
    print(json.dumps(table))


if __name__ == "__main__":
    main()
