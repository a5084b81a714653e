"""Benchmark of the M_N closure hot path on one MI355X.

Workload (the metric's configuration): 1,048,576 independent M_3 dual problems (16
moments), tensor quadrature 24 mu x 48 phi = 1152 ordinates.  Per cell s ~ U[0.5, 8],
alpha*_k ~ N(0, s^2), u = <m e^{alpha*.m}>; Newton starts from the isotropic alpha of u
(init.cpp:265) and stops at ||grad h|| <= 1e-8 (FP64) on the u0-scaled problem with
gamma <= 1.1 (optBatch.cpp:268), MN_MAX_ITER 100, MN_COND_H_MAX 1e10.  The inputs are
generated on the device once and stay in HBM; a step restores u and alpha from pristine
device copies and solves the whole batch.

Paths (--path): "fp64" plain FP64 (matrix-core Hessian sums); "mixed" the Hessian as split
FP16 matrix-core products under a log-sum-exp shift (for M_1..M_3 the column sweep: cells as
the columns of the products, Hessian as pair products) with the exponents, gradient, f,
Cholesky and the stopping test in FP64 (ill-conditioned cells continue on FP64 Hessians);
"+nn" adds the
learned warm start, trained first (not timed) on --train-pairs (u, alpha) pairs that the FP64
closure solves from a different seed.  The plain FP64 path and the reference-order FP64 path
(bit-identical to the oracle) are timed beside the chosen one (same inputs, same process);
every path's outputs are compared with the FP64 path's, and with the oracle's on the CPU
sample.  Launch shapes come from c3d-2015_amd/tuning.json (tune.py) unless --no-tuning.

One JSON line: value = cells solved per second by the chosen path; roofline of the path's
matrix-core sweep kernel (HIP events on the closure's stream, cell-sweeps counted on the
device); the CPU oracle timed on a bounded sample of the same inputs.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--path mixed+nn] [--cpu-sample 65536]
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = ("M_N cell-updates/s at ‖∇h‖≤1e-8: batched mixed+NN vs plain FP64, "
          "one GPU")
FP64_MFMA_PEAK = 78.6  # TFLOP/s, MI355X dense FP64 matrix
FP16_MFMA_PEAK = 2500.0  # TFLOP/s, MI355X dense FP16 matrix


def front_end():
    spec = importlib.util.spec_from_file_location(
        "c3d_hip", os.path.join(ROOT, "c3d-2015_amd", "c3d_hip.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["c3d_hip"] = mod
    spec.loader.exec_module(mod)
    return mod


def host_cpu_info():
    """CPU model and physical core count of this host (from /proc/cpuinfo), and the threads
    the oracle may use (OMP_NUM_THREADS, else the logical CPUs)."""
    model, cores = None, set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None and core is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or logical
    return {"cpu_model": model, "physical_cores": len(cores) or None,
            "logical_cpus": logical, "threads": threads}


def cpu_baseline(u_host, a_host, N, Q, nthreads, n_one=4096, reps=3):
    """The oracle (plain FP64 C restatement of src/opt, OpenMP schedule(dynamic) over cells)
    on a sample of the inputs: one warm-up, then the median of `reps` timed runs on all
    `nthreads` threads, and one run on a single thread over the first `n_one` cells.
    Returns (rate, seconds, rate_1thread, u_out, a_out, stats)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    T = O.Tables(N, Q)
    n = u_host.shape[0]
    O.closure(T, u_host[:256].copy(), a_host[:256].copy(), O.opts(), nthreads)
    times = []
    for _ in range(reps):
        uo, ao = u_host.copy(), a_host.copy()
        t0 = time.perf_counter()
        _, st = O.closure(T, uo, ao, O.opts(), nthreads)
        times.append(time.perf_counter() - t0)
    dt = sorted(times)[len(times) // 2]
    m = min(n_one, n)
    t0 = time.perf_counter()
    O.closure(T, u_host[:m].copy(), a_host[:m].copy(), O.opts(), 1)
    dt1 = time.perf_counter() - t0
    return n / dt, dt, m / dt1, uo, ao, st


def oracle_agreement(T, u_in, u_gpu, a_gpu, reg_gpu, flags_gpu, u_orc, a_orc, st_orc,
                     stable=None):
    """A GPU path's cells against the oracle's results on the same inputs: the share whose
    written-back moments agree to 1e-6 u0 and whose regularisation rung agrees; alpha's relative
    difference on cells of equal rung; and the FP64 stopping test of the GPU result, recomputed
    here by numpy on its unflagged cells (||<m e^{alpha.m}> - u||_2 / u0).  With `stable` (the
    cells whose oracle outcome does not change with the exp implementation), the same shares
    over those cells and the share of the disagreeing cells that are threshold cases."""
    du = np.abs(u_gpu - u_orc).max(axis=1) / np.abs(u_in[:, 0])
    same = reg_gpu == st_orc["regIndex"]
    da = np.abs(a_gpu - a_orc).max(axis=1) / np.abs(a_orc).max(axis=1)
    ok = (flags_gpu & (2 | 4 | 8)) == 0  # no u0 <= 0, ladder or budget flag
    g = ((np.exp(a_gpu[ok] @ T.M.T) * T.w) @ T.M - u_gpu[ok]) / u_gpu[ok, :1]
    extra = {}
    if stable is not None:
        agree = (du <= 1e-6) & same
        extra = {"moments_within_1e-6_stable_cells": round(float(agree[stable].mean()), 6),
                 "disagreeing_cells": int((~agree).sum()),
                 "disagreeing_threshold_cases": int((~agree & ~stable).sum())}
    return {"cells": int(len(u_in)),
            **extra,
            "moments_within_1e-6": round(float((du <= 1e-6).mean()), 6),
            "rung_agreement": round(float(same.mean()), 6),
            "alpha_max_rel_same_rung": float(da[same].max()) if same.any() else None,
            "alpha_median_rel_same_rung": float(np.median(da[same])) if same.any() else None,
            "stop_test_max_grad_norm": float(np.linalg.norm(g, axis=1).max()) if ok.any() else None}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _short_kernel(name):
    """k_name<args> from a demangled or an Itanium-mangled kernel name (rocprofv3 leaves some
    templates mangled: k_tailILi16ELb0ELb1ELi4EE... -> k_tail<16, false, true, 4>)."""
    import re
    m = re.search(r"(k_[a-z0-9_]+?)I((?:L[bi]\d+E)+)E", name)
    if m:
        args = [("true" if v == "1" else "false") if t == "b" else v
                for t, v in re.findall(r"L([bi])(\d+)E", m.group(2))]
        return f"{m.group(1)}<{', '.join(args)}>"
    m = re.search(r"(k_\w+(<[^(]*>)?|__amd_rocclr_\w+)", name)
    return m.group(1) if m else name[:48]


def mixed_sweep_work(nm, nq):
    """Algorithmic matrix work of a mixed-phase sweep per cell: the exponents alpha . m_q and
    the gradient moments sum_q w e m_q on the FP64 matrix cores (2 nq nm flops each), the
    Hessian's nm (nm + 1) / 2 distinct entries as three FP16 products (hi hi, hi lo, lo hi;
    6 nq nm (nm + 1) / 2 flops); in FP64-equivalent flops, the FP16 part weighted by the ratio
    of the two matrix peaks (time at peak rate on the FP64 matrix cores)."""
    f64 = 4.0 * nq * nm
    f16 = 6.0 * nq * nm * (nm + 1) / 2
    return {"fp64": f64, "fp16": f16, "fp64_equivalent": f64 + f16 * FP64_MFMA_PEAK / FP16_MFMA_PEAK}


def mixed_sweep_kernel(nm):
    return f"k_sweep_cols<{nm}>"


def summarize_profile(paths, steps, warmup, write_traffic=None):
    """Per-kernel figures of the timed solves of a default closure bench run, from rocprofv3
    CSV output (--output-format csv): a kernel trace (run_kernel_trace.csv: duration per
    dispatch) or a counter collection (run_counter_collection.csv: --pmc).  Dispatches are
    split into solves at each k_init launch (one per closure call): the warm start's training
    solve, then `warmup` + `steps` solves of the mixed+nn path, then the same of the FP64 path;
    the timed solves are averaged.  FETCH_SIZE / WRITE_SIZE (KiB) are calibrated on the
    timed loop's 128 MiB input copies (u.copy_(u0)); with write_traffic the bytes per dispatch
    go to that JSON file (read by the bench's roofline)."""
    import csv
    import collections
    out, traffic, calib = [], {}, {}
    for path in paths:
        for fn in sorted(os.listdir(path)):
            if not fn.endswith(".csv") or not ("kernel_trace" in fn or "counter_collection" in fn):
                continue
            rows = list(csv.DictReader(open(os.path.join(path, fn))))
            disp = collections.OrderedDict()
            for r in rows:
                d = int(r.get("Dispatch_Id") or r.get("Correlation_Id"))
                e = disp.setdefault(d, {"name": _short_kernel(r["Kernel_Name"]), "c": {}})
                e["ns"] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
                if "Counter_Name" in r:
                    e["c"][r["Counter_Name"]] = (e["c"].get(r["Counter_Name"], 0.0)
                                                 + float(r["Counter_Value"]))
            ds = [disp[k] for k in sorted(disp)]
            starts = [i for i, e in enumerate(ds) if e["name"].startswith("k_init<")]
            segs = [ds[i:(starts[j + 1] if j + 1 < len(starts) else len(ds))]
                    for j, i in enumerate(starts)]
            if len(segs) < 1 + 2 * (warmup + steps):
                raise SystemExit(f"{path}/{fn}: {len(segs)} solves, expected "
                                 f"{1 + 2 * (warmup + steps)}")
            timed = {"mixed+nn": segs[1 + warmup:1 + warmup + steps],
                     "fp64": segs[1 + 2 * warmup + steps:1 + 2 * (warmup + steps)]}
            for cn in ("FETCH_SIZE", "WRITE_SIZE"):
                big = [e["c"][cn] for e in ds if e["name"] == "__amd_rocclr_copyBuffer"
                       and e["c"].get(cn, 0) > 1e4]
                if big:
                    calib[cn] = 131072.0 / sorted(big)[len(big) // 2]
            out.append(f"### {path}/{fn}\n")
            for label, sel in timed.items():
                acc = collections.defaultdict(collections.Counter)
                cnt, ns = collections.Counter(), collections.Counter()
                for seg in sel:
                    for e in seg:
                        cnt[e["name"]] += 1
                        ns[e["name"]] += e["ns"]
                        acc[e["name"]].update(e["c"])
                cols = sorted({c for v in acc.values() for c in v})
                out.append(f"{label}, {len(sel)} timed solves (per solve):\n")
                out.append("| kernel | dispatches | ms | avg us per dispatch |"
                           + "".join(f" {c} |" for c in cols))
                out.append("|---|---|---|---|" + "---|" * len(cols))
                for k in sorted(ns, key=lambda k: -ns[k]):
                    if ns[k] / len(sel) < 1e4 and not acc[k]:
                        continue
                    out.append(f"| {k} | {cnt[k] / len(sel):.1f} | {ns[k] / len(sel) / 1e6:.3f} | "
                               f"{ns[k] / cnt[k] / 1e3:.1f} |"
                               + "".join(f" {acc[k][c] / len(sel):.4g} |" for c in cols))
                    for cn, key in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write")):
                        if cn in acc[k] and cn in calib:
                            t = traffic.setdefault(label, {}).setdefault(k, {})
                            t[key + "_bytes_per_dispatch"] = (acc[k][cn] / cnt[k] * calib[cn]
                                                              * 1024.0)
                            t["dispatches_per_solve"] = cnt[k] / len(sel)
                out.append("")
    if calib:
        out.append("Calibration on the 128 MiB input copies: " + ", ".join(
            f"{k} x {v:.4f}" for k, v in calib.items()) + " (KiB per KiB counted)")
    print("\n".join(out))
    if write_traffic and traffic:
        json.dump({"calibration": calib, "paths": traffic,
                   "command": "python3 bench.py --summarize-profile <dirs> --write-traffic ..."},
                  open(write_traffic, "w"), indent=1, sort_keys=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cells", type=int, default=1 << 20)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--seed", type=lambda s: int(s, 0), default=0xC3D2015)
    ap.add_argument("--path", choices=("fp64", "mixed", "fp64+nn", "mixed+nn"),
                    default="mixed+nn")
    ap.add_argument("--mixed-cond", type=float, default=1e8)
    ap.add_argument("--mixed-iter", type=int, default=40)
    ap.add_argument("--reforder-steps", type=int, default=2,
                    help="timed solves of the reference-order FP64 path (0: skip it)")
    ap.add_argument("--no-tuning", action="store_true",
                    help="ignore c3d-2015_amd/tuning.json (the library's default launch shapes)")
    ap.add_argument("--summarize-profile", nargs="*", default=None, metavar="DIR",
                    help="summarise rocprofv3 CSV output of a default closure bench run (run "
                         "with the same --steps/--warmup) instead of running; no GPU needed")
    ap.add_argument("--write-traffic", default=None,
                    help="with --summarize-profile: write FETCH/WRITE bytes per dispatch here")
    ap.add_argument("--async-passes", type=int, default=0,
                    help="> 0: the closure's asynchronous schedule with this many bulk passes")
    ap.add_argument("--graph", action="store_true",
                    help="step workload: capture one SSP-RK2 step in a HIP graph and time its "
                         "replays (asynchronous schedule; --async-passes, default 48); orders "
                         "workload: the four orders' closures in one graph, one launch a step")
    ap.add_argument("--tail", type=int, nargs=3, default=None, metavar=("CELLS", "ROUNDS", "WIDE"),
                    help="the closure's tail thresholds (c3d_set_tail, c3d_set_tail_wide), "
                         "overriding the tuning table")
    ap.add_argument("--finish-cond", type=int, default=None,
                    help="10 log10 of the condition indicator above which a mixed cell "
                         "finishes on FP64 Hessians (default: the library's)")
    ap.add_argument("--train-pairs", type=int, default=1 << 22)
    ap.add_argument("--train-steps", type=int, default=4000, help="MSE steps")
    ap.add_argument("--dual-steps", type=int, default=40000, help="dual-gap steps")
    ap.add_argument("--train-batch", type=int, default=16384)
    ap.add_argument("--train-lr", type=float, default=2e-3)
    ap.add_argument("--workload", choices=("closure", "step", "orders"), default="closure",
                    help="closure: the metric's batch (default); step: SSP-RK2 time "
                         "stepping (M_2, Q=8, 128^3 by default; kinetic flux roofline); "
                         "orders: 2^20 cells, a quarter each of M_1..M_4, mixed+nn vs FP64")
    ap.add_argument("--grid", type=int, default=128, help="step: cells per axis")
    ap.add_argument("--order", type=int, default=2, help="step: moment order N")
    ap.add_argument("--quad", type=int, default=8, help="step: quadrature order Q")
    ap.add_argument("--orders-streams", type=int, choices=(1, 4), default=4,
                    help="orders: 4 solves the four orders concurrently (one stream and one "
                         "host thread per order's context), 1 back to back on one stream")
    ap.add_argument("--step-precision", choices=("fp64", "mixed"), default="fp64",
                    help="step: mixed runs the mixed-precision closure's trajectory and the "
                         "reference-order FP64 one beside the FP64 one from the same state "
                         "and reports both differences from it after every step")
    args = ap.parse_args()
    if args.summarize_profile is not None:
        summarize_profile(args.summarize_profile, args.steps, args.warmup, args.write_traffic)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (closure batches split "
              "over cells with no exchange; a time-stepped grid would split into slabs with a "
              "two-cell halo exchanged each stage)", file=sys.stderr)
        sys.exit(2)

    import torch
    P = front_end()
    dev = torch.device("cuda:0")
    if args.workload == "step":
        return step_workload(args, P, torch, dev)
    if args.workload == "orders":
        return orders_workload(args, P, torch, dev)
    N, Q = 3, 24
    n = args.cells
    ctx = P.Ctx(N, Q)
    ctx.set_mixed(args.mixed_cond, args.mixed_iter)
    if args.finish_cond is not None:
        ctx.set_launch(P.LAUNCH_MIXED_FINISH_COND, args.finish_cond)
    nm, nq = ctx.nm, ctx.nq
    u0 = torch.empty((n, nm), dtype=torch.float64, device=dev)
    a0 = torch.empty_like(u0)
    ctx.make_closure_batch(n, args.seed, 0.5, 8.0, u0, a0)
    u = torch.empty_like(u0)
    a = torch.empty_like(a0)
    stats = P.alloc_stats(n, dev)
    opts = P.closure_opts(maxIter=100, tol=1e-8, tolGamma=1.1, condHMax=1e10)
    torch.cuda.synchronize()

    train = None
    if args.path.endswith("+nn"):
        log(f"training the warm start on {args.train_pairs} solver pairs")
        t0 = time.perf_counter()
        m = args.train_pairs
        ut = torch.empty((m, nm), dtype=torch.float64, device=dev)
        at = torch.empty_like(ut)
        ctx.set_precision(P.PREC_FP64)
        ctx.make_closure_batch(m, args.seed + 1, 0.5, 8.0, ut, at)
        ctx.closure(ut, at, opts)
        ctx.sync()
        t1 = time.perf_counter()
        l0, l1 = ctx.mlp_train(ut, at, mse_steps=args.train_steps, dual_steps=args.dual_steps,
                               batch=args.train_batch, lr=args.train_lr, seed=args.seed)
        t2 = time.perf_counter()
        train = {"pairs": m, "pair_seed": hex(args.seed + 1), "mse_steps": args.train_steps,
                 "dual_steps": args.dual_steps, "batch": args.train_batch, "lr": args.train_lr,
                 "mse_loss_first": round(l0, 5), "dual_gap_last": round(l1, 5), "solve_s": round(t1 - t0, 2),
                 "train_s": round(t2 - t1, 2)}
        log(f"trained: {train}")
        del ut, at
        torch.cuda.empty_cache()

    tuning = {}

    def measure(path, steps, warmup):
        prec = {"mixed": P.PREC_MIXED, "fp64": P.PREC_FP64,
                "fp64_reforder": P.PREC_FP64_REFORDER}[path.split("+")[0]]
        ctx.set_precision(prec)
        ctx.set_warmstart(path.endswith("+nn"))
        if not args.no_tuning:
            tuning[path] = ctx.apply_tuning("mixed" if path.startswith("mixed") else "fp64")
        if args.tail is not None:
            ctx.set_tail(args.tail[0], args.tail[1], wide_cells=args.tail[2])
        ctx.set_launch(P.LAUNCH_ASYNC, args.async_passes)

        def step():
            u.copy_(u0)
            a.copy_(a0)
            ctx.closure(u, a, opts, stats=stats)

        for _ in range(warmup):
            step()
        ctx.sync()
        ctx.set_profiling(True)
        acc = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
            prof = ctx.closure_profile()  # waits for the step's stream
            for k, v in prof.items():
                acc[k] = acc.get(k, 0) + v
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        rc = ctx.sync()
        ctx.set_profiling(False)
        st = P.stats_to_numpy(stats)
        out = {"ms_per_step": 1e3 * elapsed / steps, "status": rc,
               "prof": {k: v / steps for k, v in acc.items()},
               "iters_mean": float(st["totalIter"].mean()), "iters_max": int(st["totalIter"].max()),
               "iters_pct": {f"p{q}": int(np.percentile(st["totalIter"], q))
                             for q in (10, 50, 90, 99, 99.9)},
               "reg_hist": np.bincount(st["regIndex"], minlength=5).tolist(),
               "fallback": float(((st["flags"] & P.FLAG_FP64_FALLBACK) != 0).mean()),
               "warm_taken": float(((st["flags"] & P.FLAG_WARM) != 0).mean()),
               "u": u.cpu().numpy(), "a": a.cpu().numpy(), "reg": st["regIndex"].copy(),
               "flags": st["flags"].copy()}
        log(f"{path}: {out['ms_per_step']:.1f} ms/step iters {out['iters_mean']:.2f}")
        return out

    main_r = measure(args.path, args.steps, args.warmup)
    base_r = main_r if args.path == "fp64" else measure("fp64", args.steps, args.warmup)
    # the reference-order FP64 mode: the one bit-identical to the oracle (deterministic exp)
    ref_r = measure("fp64_reforder", max(1, args.reforder_steps), 0) \
        if args.reforder_steps > 0 else None
    value = n / (main_r["ms_per_step"] / 1e3)
    fp64_value = n / (base_r["ms_per_step"] / 1e3)

    # agreement of the chosen path with the FP64 path on every cell
    du = np.abs(main_r["u"] - base_r["u"]).max(axis=1) / np.abs(base_r["u"][:, 0])
    agree = {"moments_within_1e-6": round(float((du <= 1e-6).mean()), 6),
             "rung_agreement": round(float((main_r["reg"] == base_r["reg"]).mean()), 6)}

    # roofline of the path's matrix-core sweep kernel.  The FP64 path's sweep: the Hessian
    # H = M^T diag(w e) M on the FP64 matrix cores, 2 nq nm^2 flops per cell-sweep, against the
    # FP64 matrix peak.  The mixed sweep (mixed_sweep_work): FP64 exponent and gradient
    # products plus the split FP16 pair products of the Hessian, in FP64-equivalent flops
    # against the FP64 matrix peak, i.e. the fraction of the kernel's time its matrix work
    # would take at the published peaks; the rest is its FP64 vector work (exp) and stalls.

    def sweep_roofline(prof, mixed_kernel):
        cs = prof["mixed_cell_sweeps"] if mixed_kernel else prof["cell_sweeps"]
        ms = prof["mixed_sweep_ms"] if mixed_kernel else prof["sweep_ms"]
        secs = ms * 1e-3
        if mixed_kernel:
            w = mixed_sweep_work(nm, nq)
            fl = w["fp64_equivalent"]
            kern = (f"{mixed_sweep_kernel(nm)} (mixed phase: FP64 matrix-core exponents and "
                    "gradient, Hessian as split FP16 pair products)")
        else:
            fl = 2.0 * nq * nm * nm
            kern = f"k_sweep<{nm}> (FP64 MFMA Hessian)"
        ach = fl * cs / secs / 1e12 if ms > 0 else 0.0
        r = {"bound": "mfma", "achieved": round(ach, 3), "peak": FP64_MFMA_PEAK,
             "unit": "TFLOP/s", "frac": round(ach / FP64_MFMA_PEAK, 4), "traffic": None,
             "kernel": kern, "flops_per_cell_sweep": fl, "cell_sweeps_per_step": int(cs),
             "kernel_ms_per_step": round(ms, 3)}
        if mixed_kernel:
            r["basis"] = ("FP64-equivalent flops: FP64 matrix work plus FP16 matrix work x "
                          f"{FP64_MFMA_PEAK} / {FP16_MFMA_PEAK}")
            r["work_per_cell_sweep"] = {"fp64_matrix": w["fp64"], "fp16_matrix": w["fp16"]}
            r["fp64_matrix_achieved"] = round(w["fp64"] * cs / secs / 1e12, 3) if ms > 0 else 0.0
            r["fp16_matrix_achieved"] = round(w["fp16"] * cs / secs / 1e12, 3) if ms > 0 else 0.0
        return r

    # HBM bytes per dispatch of the sweep kernels, from the committed rocprofv3 PMC summary
    # (FETCH_SIZE / WRITE_SIZE runs calibrated on a copy of known size)
    # HBM bytes per dispatch of the sweep kernels in the timed solves of each path, from the
    # committed rocprofv3 PMC summary (FETCH_SIZE / WRITE_SIZE runs calibrated on the timed
    # loop's input copies)
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        for pth, ks in json.load(open(tpath)).get("paths", {}).items():
            for k, v in ks.items():
                if "fetch_bytes_per_dispatch" in v and "write_bytes_per_dispatch" in v:
                    traffic[(pth, k)] = round(v["fetch_bytes_per_dispatch"]
                                              + v["write_bytes_per_dispatch"])

    def with_traffic(r, pth):
        key = r["kernel"].split(" (")[0]
        r["traffic"] = traffic.get((pth, key))
        if r["traffic"] is not None:
            r["traffic_source"] = ("profiles/closure_C1_pmc.md (bytes per dispatch, timed "
                                   f"{pth} solves)")
        return r

    mixed_main = args.path.startswith("mixed")
    roofline = with_traffic(sweep_roofline(main_r["prof"], mixed_main), args.path)
    if mixed_main:
        roofline["fp64_sweep_in_path"] = with_traffic(sweep_roofline(main_r["prof"], False),
                                                      args.path)
    roofline["fp64_path_sweep"] = with_traffic(sweep_roofline(base_r["prof"], False), "fp64")

    def kernel_ms(r):
        p = r["prof"]
        out = {k: round(p[k], 3) for k in ("mixed_sweep_ms", "sweep_ms", "step_ms", "tail_ms",
                                            "warm_ms")}
        out.update(passes=int(p["passes"]), tail_launches=int(p["tail_launches"]))
        return out

    cpu, oracle_cmp = None, None
    if args.cpu_sample > 0:
        ns = min(args.cpu_sample, n)
        uh = u0[:ns].cpu().numpy().copy()
        ah = a0[:ns].cpu().numpy().copy()
        info = host_cpu_info()
        threads = info["threads"]
        cps, secs, cps1, uo, ao, so = cpu_baseline(uh, ah, N, Q, threads)
        cpu = {"value": round(cps, 1), "unit": "cell-updates/s", "cores": threads,
               "kind": "port",
               "sample": f"first {ns} cells of the same batch: oracle (C, -O3, OpenMP "
                         f"schedule(dynamic) over cells, plain FP64, the reference's libm exp) "
                         f"on {threads} threads, median of 3 timed runs after a warm-up: "
                         f"{secs:.2f} s; one thread on the first {min(4096, ns)} cells",
               "value_1thread": round(cps1, 1), **info}
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        T = O.Tables(N, Q)
        # threshold cases: cells whose oracle outcome (rung, moments to 1e-6 u0) changes when
        # the C library's exp is replaced by the deterministic one and the oracle's quadrature
        # and basis tables by the library's (each within ~1 ulp); the reference-order GPU path
        # is bit-identical to that second oracle run, which is checked on every sample cell
        Tl = O.Tables(N, Q)
        lt = P.tables_host(N, Q)
        for k in ("w", "xi", "eta", "mu", "M", "M2"):
            getattr(Tl, k)[...] = lt[k]  # in place: the oracle's table struct points here
        u1, a1 = uh.copy(), ah.copy()
        with O.exp_mode(1):
            _, s1 = O.closure(Tl, u1, a1, O.opts(), threads)
        stable = ((s1["regIndex"] == so["regIndex"])
                  & (np.abs(u1 - uo).max(axis=1) / np.abs(uh[:, 0]) <= 1e-6))
        oracle_cmp = {"oracle_threshold_cases": {
            "cells": int(ns), "threshold_cases": int((~stable).sum()),
            "note": "cells whose oracle rung or moments (1e-6 u0) differ between the oracle with "
                    "the C library's exp and its own tables and the oracle with the "
                    "deterministic exp and the library's tables; the GPU paths are also scored "
                    "on the other cells"}}
        for name, r in (("mixed+nn" if args.path == "mixed+nn" else args.path, main_r),
                        ("fp64", base_r)) + ((("fp64_reforder", ref_r),) if ref_r else ()):
            oracle_cmp[name] = oracle_agreement(T, uh, r["u"][:ns], r["a"][:ns], r["reg"][:ns],
                                                r["flags"][:ns], uo, ao, so, stable)
        if ref_r is not None:
            # bit for bit against the deterministic-exp oracle on the library's tables
            same = (np.all(ref_r["u"][:ns] == u1, axis=1) & np.all(ref_r["a"][:ns] == a1, axis=1)
                    & (ref_r["reg"][:ns] == s1["regIndex"]))
            oracle_cmp["fp64_reforder"]["bitwise_cells"] = int(ns)
            oracle_cmp["fp64_reforder"]["bitwise_identical_frac"] = round(float(same.mean()), 6)
        log(f"oracle agreement: {oracle_cmp}")

    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(main_r["ms_per_step"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        # the mixed phase: the Hessian as split FP16 products with FP32 accumulation; the rest
        # (exponents, f, gradient, Cholesky, stopping test) FP64
        "dtype": "f16+f32+f64" if args.path.startswith("mixed") else "f64",
        "data": "synthetic",
        "config": {"workload": "closure-only batch: 1048576 independent M_3 dual problems, "
                               "24x48=1152 ordinates, alpha*~N(0,s^2), s~U[0.5,8], "
                               "||grad h||<=1e-8 in FP64",
                   "cells": n, "moment_order": N, "quad_order": Q, "ordinates": nq,
                   "path": args.path, "status": main_r["status"],
                   "iters_per_cell_mean": round(main_r["iters_mean"], 3),
                   "iters_per_cell_max": main_r["iters_max"],
                   "iters_per_cell_percentiles": main_r["iters_pct"],
                   "regularisation_hist": main_r["reg_hist"],
                   "fp64_fallback_frac": round(main_r["fallback"], 5),
                   "warm_start_taken_frac": round(main_r["warm_taken"], 5),
                   "kernel_ms_per_step": kernel_ms(main_r),
                   "mixed_cond_max": args.mixed_cond,
                   "fp64_path": {"value": round(fp64_value, 1),
                                 "ms_per_step": round(base_r["ms_per_step"], 3),
                                 "iters_per_cell_mean": round(base_r["iters_mean"], 3),
                                 "iters_per_cell_percentiles": base_r["iters_pct"],
                                 "regularisation_hist": base_r["reg_hist"],
                                 "kernel_ms_per_step": kernel_ms(base_r)},
                   "speedup_vs_fp64": round(value / fp64_value, 3),
                   "fp64_reforder_path": None if ref_r is None else {
                       "value": round(n / (ref_r["ms_per_step"] / 1e3), 1),
                       "ms_per_step": round(ref_r["ms_per_step"], 3),
                       "iters_per_cell_mean": round(ref_r["iters_mean"], 3),
                       "regularisation_hist": ref_r["reg_hist"],
                       "speedup_of_main_path": round(ref_r["ms_per_step"]
                                                     / main_r["ms_per_step"], 3)},
                   "agreement_with_fp64": agree,
                   "agreement_with_oracle": oracle_cmp,
                   "tuning": tuning or None,
                   "closure_schedule": ("asynchronous, %d device-gated bulk passes" % args.async_passes
                                        if args.async_passes > 0 else "adaptive (host read-backs)"),
                   "warm_start_training": train},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def step_workload(args, P, torch, dev):
    """M_N (default N = 2: 9 moments), Q (default 8: 128 ordinates), periodic n^3 grid on
    [-1,1]^3, pulse rho = max(exp(-|x|^2/(2 0.05^2)), 1e-4) (1 + 0.01 U[-1,1]), sigma_s = 1,
    sigma_a = 0,
    minmod slopes (theta = 1), CFL 0.5 (dt of init.cpp:274), SSP-RK2 with the closure on
    every cell (the reference's default batched semantics, tol 1e-8), FP64.  A step is one
    full update() (update.cpp:185-275) on the device; the state keeps evolving across
    steps.  Also timed: the kinetic flux alone (ansatz + limited upwind stencil), against
    its algorithmic HBM bytes with the ansatz materialised; and P_N of the same order (the
    moment flux, linear closure) from the same initial state to the same time with its own
    dt, with its density compared to the M_N one.

    --step-precision mixed: the same problem advanced from one initial state with the FP64
    closure, with the mixed-precision one and with the reference-order FP64 one; `value` is
    the mixed trajectory's rate, and after every step the interior cells' difference from
    the FP64 trajectory is reported for the other two: max over cells of
    max_k |dalpha_k| / max_k |alpha_k|, of max_k |du_k| / u_0, of the density's relative
    difference, and the share of cells whose moments agree to 1e-6 u_0."""
    N, Q, n = args.order, args.quad, args.grid
    ctx = P.Ctx(N, Q)
    nm, nq = ctx.nm, ctx.nq
    ng = (n + 4) ** 3
    ni = n ** 3
    d = 2.0 / n
    u = torch.empty((ng, nm), dtype=torch.float64, device=dev)
    a = torch.empty_like(u)
    sT = torch.empty(ng, dtype=torch.float64, device=dev)
    sS = torch.empty_like(sT)
    ctx.init_grid(n, n, n, u, sT, sS, alpha=a, ic=1, sigma=0.05, floor=1e-4, noise=0.01,
                  seed=args.seed, cross=1.0)
    u_init = u.clone()
    theta = 1.0
    dt = 0.5 * (2.0 / (2.0 + theta)) / (3.0 / d)
    opts = P.closure_opts(maxIter=100, tol=1e-8, tolGamma=1.1, condHMax=1e10)
    stats = P.alloc_stats(2 * ng, dev)
    stream = torch.cuda.current_stream()

    def step(uu, aa):
        ctx.update(n, n, n, uu, aa, sT, sS, d, d, d, theta, dt, 1, 0, opts, stats=stats)

    mixed = args.step_precision == "mixed"
    errors, fp64_ms = [], None
    graph = None
    if args.graph or args.async_passes > 0:
        ctx.set_launch(P.LAUNCH_ASYNC, args.async_passes or 48)
    if not mixed:
        for _ in range(max(1, args.warmup)):
            step(u, a)
        rc = ctx.sync()
        if args.graph:
            # the eager steps sized the context's buffers; the capture records one step (the
            # captured work does not run) and each replay advances the state by one step
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph):
                step(u, a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            if graph is not None:
                graph.replay()
            else:
                step(u, a)
        rc = rc or ctx.sync()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / args.steps
    else:
        # three trajectories from the same state: the FP64 closure's, the mixed one's and,
        # as the yardstick of the FP64 closure's own sensitivity to rounding, the
        # reference-order FP64 one (other summation order, other exp); each step is timed
        # on its own (the other trajectories' steps are not in it)
        uF, aF = u.clone(), a.clone()
        uR, aR = u.clone(), a.clone()
        rc, times = 0, {"fp64": [], "mixed": [], "reforder": []}
        inner = lambda x: x.view(n + 4, n + 4, n + 4, nm)[2:-2, 2:-2, 2:-2].reshape(-1, nm)

        def diff(uu, aa):
            ui, uf, ai, af = inner(uu), inner(uF), inner(aa), inner(aF)
            du = (ui - uf).abs().amax(dim=1) / uf[:, 0].abs()
            da = (ai - af).abs().amax(dim=1) / af.abs().amax(dim=1)
            drho = (ui[:, 0] - uf[:, 0]).abs() / uf[:, 0].abs()
            return {"alpha_max_rel": float(da.max()), "moments_max_rel_u0": float(du.max()),
                    "rho_max_rel": float(drho.max()),
                    "cells_moments_within_1e-6": round(float((du <= 1e-6).double().mean()), 6)}

        for k in range(args.warmup + args.steps):
            for name, (uu, aa), prec in (("fp64", (uF, aF), P.PREC_FP64),
                                         ("mixed", (u, a), P.PREC_MIXED),
                                         ("reforder", (uR, aR), P.PREC_FP64_REFORDER)):
                ctx.set_precision(prec)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                step(uu, aa)
                torch.cuda.synchronize()
                if k >= args.warmup:
                    times[name].append(time.perf_counter() - t0)
                rc = rc or ctx.sync()
            e = {"step": k + 1, "mixed": diff(u, a), "fp64_reference_order": diff(uR, aR)}
            errors.append(e)
            log(f"step {k + 1}: " + " | ".join(
                f"{key}: alpha {v['alpha_max_rel']:.2e} moments {v['moments_max_rel_u0']:.2e} "
                f"rho {v['rho_max_rel']:.2e} within 1e-6 {v['cells_moments_within_1e-6']}"
                for key, v in e.items() if key != "step"))
        ctx.set_precision(P.PREC_FP64)
        ms = 1e3 * sum(times["mixed"]) / args.steps
        fp64_ms = 1e3 * sum(times["fp64"]) / args.steps
    st = P.stats_to_numpy(stats)
    # P_N of the same order on the same initial state (flux_moment.cpp, linear closure), with
    # its own dt (init.cpp:237) and main.cpp's clipped last step, up to the time the M_N run
    # reached: its cost per step and its density against the M_N one at equal t
    pn = None
    if not mixed:
        t_end = (args.warmup + args.steps) * dt
        dt_pn = 0.5 * 0.5 * d
        up = u_init.clone()
        t_pn, k_pn = 0.0, 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while t_pn < t_end - 1e-14 * t_end:
            h = min(dt_pn, t_end - t_pn)
            ctx.update(n, n, n, up, None, sT, sS, d, d, d, theta, h, 0, 0, opts)
            t_pn += h
            k_pn += 1
        torch.cuda.synchronize()
        pn_ms = 1e3 * (time.perf_counter() - t0) / k_pn
        rm = u.view(n + 4, n + 4, n + 4, nm)[2:-2, 2:-2, 2:-2, 0]
        rp = up.view(n + 4, n + 4, n + 4, nm)[2:-2, 2:-2, 2:-2, 0]
        pn = {"ms_per_step": round(pn_ms, 3), "value": round(ni / (pn_ms / 1e3), 1),
              "steps": k_pn, "t": t_end, "dt": dt_pn,
              "rho_rel_l2_diff_vs_mn": float(torch.linalg.vector_norm(rp - rm)
                                             / torch.linalg.vector_norm(rm)),
              "rho_max_rel_diff_vs_mn": float(((rp - rm).abs() / rm.abs()).max()),
              "min_rho_pn": float(rp.min() / (2.0 * math.pi ** 0.5)),
              "min_rho_mn": float(rm.min() / (2.0 * math.pi ** 0.5))}
        del up
    # the kinetic flux alone, HIP events on the stream it is launched on
    flux = torch.zeros_like(u)
    reps = 5
    ctx.flux_kinetic(n, n, n, a, flux, d, d, d, theta, stream=stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        ctx.flux_kinetic(n, n, n, a, flux, d, d, d, theta, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    flux_ms = e0.elapsed_time(e1) / reps
    flux_bytes = 8.0 * (ng * nm + 2.0 * ng * nq + ni * nm)
    gbs = flux_bytes / (flux_ms * 1e-3) / 1e9
    # the same call against the FP64 vector peak: ansatz exponents (2 nm + 1 per cell and
    # ordinate on every cell) and the limited stencil (~90 + 2 nm per interior cell and ordinate)
    flux_flops = ng * nq * (2.0 * nm + 1.0) + ni * nq * (90.0 + 2.0 * nm)
    tfs = flux_flops / (flux_ms * 1e-3) / 1e12
    traffic = None  # PMC-measured bytes per flux call, profiled for M_2, Q = 8, 128^3 only
    tp = os.path.join(ROOT, "profiles", "pmc_traffic_step.json")
    if os.path.exists(tp) and (N, Q, n) == (2, 8, 128):
        traffic = json.load(open(tp)).get("flux_call_bytes")
    cpu = None
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        T = O.Tables(N, Q)
        m = 32 if nm * nq <= 2048 else 12
        dm = 2.0 / m
        uc, sTc, sSc = O.init_grid(m, m, m, nm, ic=1, sigma=0.05, floor=1e-4, noise=0.01,
                                   seed=args.seed)
        ac = O.init_alpha(uc)
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        ngc = uc.shape[0]
        dtc = 0.5 * (2.0 / 3.0) / (3.0 / dm)
        t0 = time.perf_counter()
        O.lib().orc_update(m, m, m, nm, T.nm2, nq, 1, 0, uc, ac, np.zeros_like(uc),
                           np.zeros_like(uc), np.zeros(ngc * nq), sTc, sSc, O.C.byref(T.c),
                           T.ops, O.C.byref(O.opts()), dm, dm, dm, theta, dtc, None, threads)
        secs = time.perf_counter() - t0
        cpu = {"value": round(m ** 3 / secs, 1), "unit": "cell-updates/s", "cores": threads,
               "kind": "port", "sample": f"one SSP-RK2 step of the oracle on the same "
                                         f"problem at {m}^3, {secs:.1f} s"}
    prec = ("mixed-precision closure (split FP16 matrix-core Hessian, FP64 exponents, gradient "
            "and stopping test)" if mixed
            else "FP64 closure")
    out = {
        "metric": f"M_{N} SSP-RK2 interior cell-updates/s (time stepping, one GPU)",
        "value": round(ni / (ms / 1e3), 1), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16+f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"M_{N} time stepping, {n}^3 periodic grid, Q={Q} ({nq} "
                               "ordinates), pulse sigma 0.05 + 1% noise, theta 1, CFL 0.5, "
                               f"closure tol 1e-8, {prec}",
                   "step_launch": ("one captured HIP graph per step, replayed" if graph is not None
                                   else "eager c3d_update calls"),
                   "closure_schedule": ("asynchronous" if args.graph or args.async_passes > 0
                                        else "adaptive"),
                   "grid": n, "moment_order": N, "quad_order": Q, "dt": dt, "status": rc,
                   "closure_iters_mean": round(float(st["totalIter"].mean()), 3),
                   "closure_iters_max": int(st["totalIter"].max())},
        "pn_same_grid": pn,
        "fp64_trajectory": None if not mixed else {
            "ms_per_step": round(fp64_ms, 3), "value": round(ni / (fp64_ms / 1e3), 1),
            "speedup_of_mixed": round(fp64_ms / ms, 3),
            "reference_order_ms_per_step": round(1e3 * sum(times["reforder"]) / args.steps, 3),
            "difference_from_fp64_after_each_step": errors},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(gbs / 8000.0, 4),
                     "traffic": round(traffic) if traffic else None,
                     "traffic_source": "profiles/step_C2_flux_pmc_traffic.md" if traffic else None,
                     "kernel": "kinetic flux (k_ansatz + k_stencil), ansatz materialised",
                     "bytes_per_launch": flux_bytes, "ms_per_launch": round(flux_ms, 4),
                     "fp64_valu": {"achieved": round(tfs, 2), "peak": 78.6, "unit": "TFLOP/s",
                                   "frac": round(tfs / 78.6, 4),
                                   "flops_per_launch": flux_flops}},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def orders_workload(args, P, torch, dev):
    """2^20 independent cells, a quarter each of M_1, M_2, M_3, M_4 (Q = 24), alpha* as in the
    metric's batch.  Each order has its own context and warm-start network (trained on a
    quarter of --train-pairs solver pairs of that order); one step solves the four quarters,
    mixed+nn, then the same with the plain FP64 path, each order with its tuned launch shapes.
    With --orders-streams 4 (default) the four solves run concurrently, each on its own
    stream, so one order's latency-bound tail overlaps the others' bulk passes: with the
    asynchronous closure schedule (--async-passes > 0) one host thread queues all four (a call
    returns once its work is queued); otherwise each stream has a host thread of its own (the
    adaptive schedule waits on its stream between passes).  With 1 they run back to back on
    one stream."""
    Q, n = 24, args.cells
    opts = P.closure_opts(maxIter=100, tol=1e-8, tolGamma=1.1, condHMax=1e10)
    parts = []
    for k, N in enumerate((1, 2, 3, 4)):
        ctx = P.Ctx(N, Q)
        ctx.set_mixed(args.mixed_cond, args.mixed_iter)
        nm = ctx.nm
        nc = n // 4
        u0 = torch.empty((nc, nm), dtype=torch.float64, device=dev)
        a0 = torch.empty_like(u0)
        ctx.make_closure_batch(nc, args.seed + 16 * k, 0.5, 8.0, u0, a0)
        m = args.train_pairs // 4
        ut = torch.empty((m, nm), dtype=torch.float64, device=dev)
        at = torch.empty_like(ut)
        ctx.set_precision(P.PREC_FP64)
        ctx.make_closure_batch(m, args.seed + 16 * k + 1, 0.5, 8.0, ut, at)
        ctx.closure(ut, at, opts)
        ctx.sync()
        ctx.mlp_train(ut, at, mse_steps=args.train_steps, dual_steps=args.dual_steps,
                      batch=args.train_batch, lr=args.train_lr, seed=args.seed)
        del ut, at
        parts.append({"N": N, "ctx": ctx, "u0": u0, "a0": a0, "u": torch.empty_like(u0),
                      "a": torch.empty_like(a0), "st": P.alloc_stats(nc, dev)})
    torch.cuda.empty_cache()
    log("trained four warm starts")

    concurrent = args.orders_streams > 1
    if args.graph and args.async_passes == 0:
        args.async_passes = 48
    for p in parts:
        p["stream"] = torch.cuda.Stream() if concurrent or args.graph else torch.cuda.current_stream()
        p["ctx"].set_launch(P.LAUNCH_ASYNC, args.async_passes)
    pool = (ThreadPoolExecutor(max_workers=len(parts))
            if concurrent and args.async_passes == 0 else None)

    def solve(p):
        with torch.cuda.stream(p["stream"]):
            p["u"].copy_(p["u0"])
            p["a"].copy_(p["a0"])
            p["ctx"].closure(p["u"], p["a"], opts, stats=p["st"], stream=p["stream"])

    def run(path, steps, warmup):
        for p in parts:
            p["ctx"].set_precision(P.PREC_MIXED if path == "mixed+nn" else P.PREC_FP64)
            p["ctx"].set_warmstart(path == "mixed+nn")
            p["tune"] = None
            if not args.no_tuning:
                p["tune"] = p["ctx"].apply_tuning("mixed" if path == "mixed+nn" else "fp64")

        def step():
            if pool is None:
                for p in parts:
                    solve(p)
            else:
                for f in [pool.submit(solve, p) for p in parts]:
                    f.result()

        for _ in range(max(1, warmup)):
            step()
        for p in parts:
            p["ctx"].sync(p["stream"])

        def profiled(nsteps):
            # HIP-event spans of each order's kernels on its stream (waits for the stream)
            for p in parts:
                p["ctx"].set_profiling(True)
                p["prof"] = {}
            for _ in range(nsteps):
                step()
                for p in parts:
                    for k, v in p["ctx"].closure_profile().items():
                        p["prof"][k] = p["prof"].get(k, 0) + v / nsteps
            for p in parts:
                p["ctx"].set_profiling(False)

        graph = None
        if args.graph:
            # the four closures in one HIP graph: forked from the capture stream onto the
            # orders' streams and joined back, so one graph launch solves the whole batch
            graph, main = torch.cuda.CUDAGraph(), torch.cuda.Stream()
            with torch.cuda.graph(graph, stream=main):
                fork = torch.cuda.Event()
                fork.record(main)
                for p in parts:
                    p["stream"].wait_event(fork)
                    solve(p)
                for p in parts:
                    join = torch.cuda.Event()
                    join.record(p["stream"])
                    main.wait_event(join)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if graph is not None:
            for _ in range(steps):
                graph.replay()
        else:
            profiled(steps)
        rc = 0
        for p in parts:
            rc = rc or p["ctx"].sync(p["stream"])
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        if graph is not None:
            profiled(1)  # kernel figures from one eager solve after the timed replays
            for p in parts:
                p["ctx"].sync(p["stream"])
        per = []
        for p in parts:
            p["ctx"].set_profiling(False)
            st = P.stats_to_numpy(p["st"])
            pr = p["prof"]
            mixed_k = path == "mixed+nn"
            sweeps = pr["mixed_cell_sweeps"] if mixed_k else pr["cell_sweeps"]
            kms = pr["mixed_sweep_ms"] if mixed_k else pr["sweep_ms"]
            nm_, nq_ = p["ctx"].nm, p["ctx"].nq
            per.append({"N": p["N"], "iters_mean": round(float(st["totalIter"].mean()), 3),
                        "sweep_kernel": (mixed_sweep_kernel(nm_) if mixed_k
                                         else f"k_sweep<{nm_}>"),
                        "sweep_ms": round(kms, 3), "cell_sweeps": int(sweeps),
                        "flops_per_cell_sweep": (mixed_sweep_work(nm_, nq_)["fp64_equivalent"]
                                                 if mixed_k else 2 * nq_ * nm_ * nm_),
                        "fallback_frac": round(float(((st["flags"] & P.FLAG_FP64_FALLBACK) != 0)
                                                     .mean()), 5),
                        "u": p["u"].cpu().numpy(), "reg": st["regIndex"].copy()})
        log(f"{path}: {ms:.1f} ms/step")
        return ms, rc, per

    ms_m, rc_m, per_m = run("mixed+nn", args.steps, args.warmup)
    ms_f, rc_f, per_f = run("fp64", args.steps, args.warmup)
    # roofline: the order whose sweep kernel takes the longest, on its matrix-core peak
    dom = max(per_m, key=lambda r: r["sweep_ms"])
    ach = dom["flops_per_cell_sweep"] * dom["cell_sweeps"] / (dom["sweep_ms"] / 1e3) / 1e12
    roof = {"bound": "mfma", "achieved": round(ach, 3), "peak": FP64_MFMA_PEAK,
            "unit": "TFLOP/s", "frac": round(ach / FP64_MFMA_PEAK, 4), "traffic": None,
            "kernel": dom["sweep_kernel"] + f" (M_{dom['N']} quarter; FP64 exponents and "
                      "gradient, split FP16 Hessian, in FP64-equivalent flops as the closure "
                      "line counts them; HIP-event spans overlap other orders' kernels)",
            "per_order": [{k: r[k] for k in ("N", "sweep_kernel", "sweep_ms", "cell_sweeps",
                                                "flops_per_cell_sweep")} for r in per_m]}
    # CPU baseline: the oracle on the first cells of each quarter
    cpu = None
    if args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        info = host_cpu_info()
        m = max(1, min(args.cpu_sample // 4, n // 4))
        tot = 0.0
        for p in parts:
            T = O.Tables(p["N"], Q)
            uh = p["u0"][:m].cpu().numpy().copy()
            ah = p["a0"][:m].cpu().numpy().copy()
            t0 = time.perf_counter()
            O.closure(T, uh, ah, O.opts(), info["threads"])
            tot += time.perf_counter() - t0
        cpu = {"value": round(4 * m / tot, 1), "unit": "cell-updates/s",
               "cores": info["threads"], "kind": "port",
               "sample": f"first {m} cells of each order's quarter: oracle (C, -O3, OpenMP "
                         f"schedule(dynamic), plain FP64) on {info['threads']} threads: "
                         f"{tot:.2f} s", **info}
    orders = []
    for pm, pf in zip(per_m, per_f):
        du = np.abs(pm["u"] - pf["u"]).max(axis=1) / np.abs(pf["u"][:, 0])
        orders.append({"N": pm["N"], "iters_mixed_nn": pm["iters_mean"],
                       "iters_fp64": pf["iters_mean"], "fallback_frac": pm["fallback_frac"],
                       "moments_within_1e-6": round(float((du <= 1e-6).mean()), 6),
                       "rung_agreement": round(float((pm["reg"] == pf["reg"]).mean()), 6)})
    out = {
        "metric": "M_N cell-updates/s at ||grad h||<=1e-8, orders 1-4 mixed, mixed+NN, one GPU",
        "value": round(n / (ms_m / 1e3), 1), "unit": "cell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_m, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16+f32+f64", "data": "synthetic",
        "config": {"workload": f"closure-only batch: {n} cells, a quarter each of M_1..M_4, "
                               "Q=24, alpha*~N(0,s^2), s~U[0.5,8]",
                   "orders_run": ("one HIP graph launch per step: the four closures forked onto "
                                  f"four streams (asynchronous schedule, {args.async_passes} bulk "
                                  "passes)" if args.graph else
                                  "back to back on one stream" if not concurrent else
                                  "concurrently, one stream per order, queued by one host "
                                  f"thread (asynchronous schedule, {args.async_passes} bulk "
                                  "passes)" if args.async_passes > 0 else
                                  "concurrently, one stream and host thread per order"),
                   "status": rc_m, "fp64_ms_per_step": round(ms_f, 3), "fp64_status": rc_f,
                   "speedup_vs_fp64": round(ms_f / ms_m, 3), "orders": orders},
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
