"""Python front end of libc3d_hip.so (ctypes over the C-ABI in include/c3d_hip.h).

Device buffers are passed as raw device pointers; torch tensors on a ROCm device are a
convenient way to own them (see Ctx methods, which accept tensors or ints).  The library
must be built (``make -C c3d-2015_amd``); loading fails loudly otherwise, and nothing here
falls back to a CPU path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libc3d_hip.so")

(C3D_OK, C3D_ERR_ARG, C3D_ERR_HIP, C3D_ERR_U0, C3D_ERR_LADDER, C3D_ERR_UNSUPPORTED,
 C3D_ERR_BUDGET, C3D_ERR_SCRATCH) = range(8)
NOALLOC_SCRATCH_BYTES = 128 << 20
FLAG_ISO_RESTART, FLAG_U0_NONPOS, FLAG_LADDER, FLAG_BUDGET, FLAG_FP64_FALLBACK = 1, 2, 4, 8, 16
PREC_FP64, PREC_FP64_REFORDER, PREC_MIXED = 0, 1, 2
LAUNCH_MIXED_WAVES, LAUNCH_MIXED_FINISH_COND, LAUNCH_ASYNC = 1, 2, 3

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_d = C.c_double


class ClosureProfile(C.Structure):
    _fields_ = [("cell_sweeps", _i64), ("tail_cell_sweeps", _i64), ("mixed_cell_sweeps", _i64),
                ("passes", _i32), ("tail_launches", _i32), ("sweep_ms", _d),
                ("mixed_sweep_ms", _d), ("step_ms", _d), ("tail_ms", _d), ("warm_ms", _d)]


MLP_IN, MLP_HID, MLP_NPARAM = 25, 128, 23065
FLAG_WARM = 32
FLAG_FP64_FINISH = 64


class MlpNorm(C.Structure):
    _fields_ = [("x_mean", C.c_float * MLP_IN), ("x_std", C.c_float * MLP_IN),
                ("alpha_mean", C.c_float * MLP_IN), ("alpha_std", C.c_float * MLP_IN)]

    def to_numpy(self):
        import numpy as np
        return {k: np.array(getattr(self, k)[:], dtype=np.float32) for k, _ in self._fields_}

    @classmethod
    def from_numpy(cls, d):
        n = cls()
        for k, _ in cls._fields_:
            getattr(n, k)[:] = [float(x) for x in d[k]]
        return n


class ClosureOpts(C.Structure):
    _fields_ = [("maxIter", _i32), ("useGaunt", _i32), ("useGauntSparse", _i32),
                ("unbatched", _i32), ("tol", _d), ("tolGamma", _d), ("condHMax", _d)]


def closure_opts(maxIter=100, tol=1e-8, tolGamma=1.1, condHMax=1e10, useGaunt=False,
                 useGauntSparse=False, unbatched=False):
    return ClosureOpts(maxIter, int(useGaunt), int(useGauntSparse), int(unbatched), tol,
                       tolGamma, condHMax)


# (name, restype, argtypes) for every symbol of include/c3d_hip.h
SIGNATURES = [
    ("solveFluxKinetic_hip", None, [C.c_int] * 5 + [_d] * 3 + [_vp] * 7 + [C.c_int]),
    ("solveHInit_hip", None, [C.c_int] * 4 + [_vp] * 7 + [C.c_int]),
    ("solveH_hip", None, [C.c_int] * 4 + [_vp, _vp, C.c_int, C.c_int]),
    ("c3d_seam_status", C.c_int, []),
    ("c3d_create", C.c_int, [C.POINTER(_vp), C.c_int, C.c_int]),
    ("c3d_create_from_tables", C.c_int, [C.POINTER(_vp), C.c_int, C.c_int] + [_vp] * 7 +
     [C.c_int, _vp, _vp, _vp]),
    ("c3d_destroy", C.c_int, [_vp]),
    ("c3d_tables_host", C.c_int, [C.c_int, C.c_int] + [_vp] * 9 + [C.c_int, _vp, _vp, _vp]),
    ("c3d_sizes", C.c_int, [_vp] + [C.POINTER(C.c_int)] * 4),
    ("c3d_get_tables", C.c_int, [_vp] * 12),
    ("c3d_closure", C.c_int, [_vp, _i64, _vp, _vp, C.POINTER(ClosureOpts), _vp, _vp]),
    ("c3d_closure_noalloc", C.c_int, [_vp, _i64, _vp, _vp, C.POINTER(ClosureOpts), _vp,
                                      _vp]),
    ("c3d_set_profiling", C.c_int, [_vp, C.c_int]),
    ("c3d_get_closure_profile", C.c_int, [_vp, C.POINTER(ClosureProfile)]),
    ("c3d_set_tail", C.c_int, [_vp, _i32, _i32]),
    ("c3d_set_tail_wide", C.c_int, [_vp, _i32]),
    ("c3d_set_launch", C.c_int, [_vp, C.c_int, C.c_int]),
    ("c3d_set_precision", C.c_int, [_vp, C.c_int]),
    ("c3d_set_mixed", C.c_int, [_vp, _d, C.c_int, _d]),
    ("c3d_set_pass_budget", C.c_int, [_vp, C.c_int]),
    ("c3d_hessian", C.c_int, [_vp, _i64, _vp, _vp, C.c_int, C.c_int, _vp]),
    ("c3d_dense_solve", C.c_int, [C.c_int, _i64] + [_vp] * 6 + [C.c_int, _vp]),
    ("c3d_sync", C.c_int, [_vp, _vp]),
    ("c3d_ghost_fill", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    ("c3d_flux_kinetic", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _d, _d, _d, _d,
                                   C.c_int, _vp]),
    ("c3d_flux_moment", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _d, _d, _d, _vp]),
    ("c3d_euler", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _d, _vp]),
    ("c3d_rk2_combine", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int, _vp]),
    ("c3d_update", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _d, _d, _d,
                             _d, _d, C.c_int, C.c_int, C.POINTER(ClosureOpts), _vp, _vp]),
    ("c3d_make_closure_batch", C.c_int, [_vp, _i64, C.c_uint64, _d, _d, _vp, _vp, _vp, _vp]),
    ("c3d_stats_summary", C.c_int, [_vp, _i64, _vp, _vp, _vp]),
    ("c3d_init_grid", C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int, _d, _d, _d,
                                C.c_uint64, _d, _vp, _vp, _vp, _vp, _vp]),
    ("c3d_mlp_train", C.c_int, [_vp, _i64, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float,
                                C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_float), _vp]),
    ("c3d_mlp_dual_loss_grad", C.c_int, [_vp, _i64, _vp, _vp, _vp, C.POINTER(MlpNorm), _vp, _vp,
                                         _vp]),
    ("c3d_mlp_get", C.c_int, [_vp, _vp, C.POINTER(MlpNorm)]),
    ("c3d_mlp_set", C.c_int, [_vp, _vp, C.POINTER(MlpNorm)]),
    ("c3d_set_warmstart", C.c_int, [_vp, C.c_int]),
    ("c3d_mlp_features", C.c_int, [_vp, _i64, _vp, _vp, C.POINTER(MlpNorm), _vp, _vp, _vp]),
    ("c3d_mlp_forward", C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp]),
    ("c3d_mlp_infer", C.c_int, [_vp, _i64, _vp, _vp, _vp]),
    ("c3d_mlp_loss_grad", C.c_int, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
]

_lib = None


def load():
    """Load libc3d_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C c3d-2015_amd`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(x):
    """device pointer of a tensor / int / None"""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream(s):
    if s is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(s, int):
        return s
    return s.cuda_stream


class StatsSum(C.Structure):
    _fields_ = [("cells", C.c_int64), ("sumIter", C.c_int64), ("sumIterGamma", C.c_int64),
                ("maxIter", C.c_int64), ("maxIterGamma", C.c_int64), ("rung", C.c_int64 * 5)]


class C3DError(RuntimeError):
    def __init__(self, fn, rc):
        super().__init__(f"{fn} returned {rc}")
        self.rc = rc


def check(fn, rc):
    if rc != C3D_OK:
        raise C3DError(fn, rc)
    return rc


TUNING_TABLE = os.path.join(HERE, "tuning.json")


def tuning_for(N, Q, precision, path=None):
    """The tuned launch shapes of the closure for (N, Q, "fp64" | "mixed") from the table that
    tune.py writes (None if the table or the entry is absent)."""
    import json
    path = path or TUNING_TABLE
    if not os.path.exists(path):
        return None
    return json.load(open(path)).get(f"{N},{Q},{precision}")


class Ctx:
    """A c3d context (tables resident on the current device)."""

    def __init__(self, N=None, Q=None, tables=None):
        L = load()
        h = _vp()
        if tables is None:
            check("c3d_create", L.c3d_create(C.byref(h), N, Q))
        else:
            T = tables
            nnz = int(T.nnz)
            check("c3d_create_from_tables", L.c3d_create_from_tables(
                C.byref(h), T.N, T.Q, T.w.ctypes.data, T.xi.ctypes.data, T.eta.ctypes.data,
                T.mu.ctypes.data, T.M.ctypes.data, T.M2.ctypes.data, T.beta.ctypes.data, nnz,
                T.betaVal.ctypes.data if nnz else None,
                T.betaColInd.ctypes.data if nnz else None,
                T.betaRowPtr.ctypes.data if nnz else None))
        self.h = h
        nm, nm2, nq, nnz = (C.c_int() for _ in range(4))
        L.c3d_sizes(h, C.byref(nm), C.byref(nm2), C.byref(nq), C.byref(nnz))
        self.nm, self.nm2, self.nq, self.nnz = nm.value, nm2.value, nq.value, nnz.value
        self.L = L

    def close(self):
        if self.h:
            self.L.c3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def closure(self, u, alpha, opts=None, stats=None, stream=None, ncells=None):
        if opts is None:
            opts = closure_opts()
        n = ncells if ncells is not None else u.numel() // self.nm
        return check("c3d_closure", self.L.c3d_closure(self.h, n, _p(u), _p(alpha),
                                                       C.byref(opts), _p(stats),
                                                       _stream(stream)))

    def set_profiling(self, on):
        return check("c3d_set_profiling", self.L.c3d_set_profiling(self.h, int(on)))

    def closure_profile(self):
        p = ClosureProfile()
        check("c3d_get_closure_profile", self.L.c3d_get_closure_profile(self.h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in ClosureProfile._fields_}

    def set_precision(self, mode):
        return check("c3d_set_precision", self.L.c3d_set_precision(self.h, mode))

    def set_mixed(self, cond_max=1e8, max_iter=40, switch_tol=5e-4):
        return check("c3d_set_mixed", self.L.c3d_set_mixed(self.h, cond_max, max_iter,
                                                           switch_tol))

    def apply_tuning(self, precision, path=None):
        """Set the closure's launch shapes from the tuning table for this context's (N, Q) and
        `precision` ("fp64" | "mixed"); returns the entry applied, or None."""
        import math
        N = int(round(math.sqrt(self.nm))) - 1
        Q = int(round(math.sqrt(self.nq / 2)))
        t = tuning_for(N, Q, precision, path)
        if t is None:
            return None
        self.set_tail(t["tail_cells"], t["tail_passes"], wide_cells=t["tail_wide"])
        return t

    def set_launch(self, param, value):
        return check("c3d_set_launch", self.L.c3d_set_launch(self.h, param, int(value)))

    def set_pass_budget(self, n):
        return check("c3d_set_pass_budget", self.L.c3d_set_pass_budget(self.h, n))

    def set_tail(self, max_cells, passes_per_launch=32, wide_cells=None):
        check("c3d_set_tail", self.L.c3d_set_tail(self.h, max_cells, passes_per_launch))
        if wide_cells is not None:
            check("c3d_set_tail_wide", self.L.c3d_set_tail_wide(self.h, wide_cells))

    def hessian(self, alpha, hs, stream=None, gaunt=False, sparse=False):
        n = alpha.numel() // self.nm
        return check("c3d_hessian", self.L.c3d_hessian(self.h, n, _p(alpha), _p(hs), int(gaunt),
                                                       int(sparse), _stream(stream)))

    def sync(self, stream=None):
        return self.L.c3d_sync(self.h, _stream(stream))

    def stats_summary(self, stats, ncells=None, stream=None):
        """Sums of per-cell outcomes on the device (c3d_stats_summary) as a dict."""
        n = ncells if ncells is not None else stats.numel() * stats.element_size() // STATS_BYTES
        out = StatsSum()
        check("c3d_stats_summary", self.L.c3d_stats_summary(self.h, n, _p(stats), C.byref(out),
                                                             _stream(stream)))
        d = {k: getattr(out, k) for k, _ in StatsSum._fields_ if k != "rung"}
        d["rung"] = list(out.rung)
        return d

    def ghost_fill(self, nx, ny, nz, u, stream=None):
        return check("c3d_ghost_fill", self.L.c3d_ghost_fill(self.h, nx, ny, nz, _p(u),
                                                             _stream(stream)))

    def flux_kinetic(self, nx, ny, nz, alpha, flux, dx, dy, dz, theta, isPn=0, stream=None):
        return check("c3d_flux_kinetic", self.L.c3d_flux_kinetic(
            self.h, nx, ny, nz, _p(alpha), _p(flux), dx, dy, dz, theta, int(isPn),
            _stream(stream)))

    def flux_moment(self, nx, ny, nz, u, flux, dx, dy, dz, stream=None):
        return check("c3d_flux_moment", self.L.c3d_flux_moment(self.h, nx, ny, nz, _p(u),
                                                               _p(flux), dx, dy, dz,
                                                               _stream(stream)))

    def euler(self, nx, ny, nz, u, flux, sigmaT, sigmaS, dt, stream=None):
        return check("c3d_euler", self.L.c3d_euler(self.h, nx, ny, nz, _p(u), _p(flux),
                                                   _p(sigmaT), _p(sigmaS), dt, _stream(stream)))

    def rk2_combine(self, nx, ny, nz, u, uOld, kind, stream=None):
        return check("c3d_rk2_combine", self.L.c3d_rk2_combine(self.h, nx, ny, nz, _p(u),
                                                               _p(uOld), kind, _stream(stream)))

    def update(self, nx, ny, nz, u, alpha, sigmaT, sigmaS, dx, dy, dz, theta, dt, solver=1,
               pnUseKinetic=0, opts=None, stats=None, stream=None):
        if opts is None:
            opts = closure_opts()
        return check("c3d_update", self.L.c3d_update(
            self.h, nx, ny, nz, _p(u), _p(alpha), _p(sigmaT), _p(sigmaS), dx, dy, dz, theta, dt,
            solver, pnUseKinetic, C.byref(opts), _p(stats), _stream(stream)))

    def make_closure_batch(self, n, seed, smin, smax, u, alpha0, alphaStar=None, stream=None):
        return check("c3d_make_closure_batch", self.L.c3d_make_closure_batch(
            self.h, n, seed, smin, smax, _p(u), _p(alpha0), _p(alphaStar), _stream(stream)))

    def init_grid(self, nx, ny, nz, u, sigmaT, sigmaS, alpha=None, a=(-1.0, -1.0, -1.0),
                  d=None, ic=1, sigma=0.1, floor=1e-4, noise=0.0, seed=0, cross=1.0,
                  stream=None):
        import numpy as np
        if d is None:
            d = ((1.0 - a[0]) / nx, (1.0 - a[1]) / ny, (1.0 - a[2]) / nz)
        av = np.ascontiguousarray(a, dtype=np.float64)
        dv = np.ascontiguousarray(d, dtype=np.float64)
        return check("c3d_init_grid", self.L.c3d_init_grid(
            self.h, nx, ny, nz, av.ctypes.data, dv.ctypes.data, ic, sigma, floor, noise, seed,
            cross, _p(u), _p(alpha), _p(sigmaT), _p(sigmaS), _stream(stream)))

    # ---- learned warm start
    def mlp_train(self, u, alpha, mse_steps, dual_steps=0, batch=4096, lr=2e-3, seed=1,
                  stream=None):
        """Train the context's network on (u, alpha) device pairs; returns (first, last) loss."""
        n = u.numel() // self.nm
        l0, l1 = C.c_float(), C.c_float()
        check("c3d_mlp_train", self.L.c3d_mlp_train(self.h, n, _p(u), _p(alpha), mse_steps,
                                                    dual_steps, batch, lr, seed, C.byref(l0),
                                                    C.byref(l1), _stream(stream)))
        return l0.value, l1.value

    def mlp_dual_loss_grad(self, params, X, u, norm, grad, loss, stream=None):
        n = X.numel() // MLP_IN
        if isinstance(norm, dict):
            norm = MlpNorm.from_numpy(norm)
        return check("c3d_mlp_dual_loss_grad", self.L.c3d_mlp_dual_loss_grad(
            self.h, n, _p(params), _p(X), _p(u), C.byref(norm), _p(grad), _p(loss),
            _stream(stream)))

    def mlp_get(self):
        import numpy as np
        p = np.zeros(MLP_NPARAM, dtype=np.float32)
        nrm = MlpNorm()
        check("c3d_mlp_get", self.L.c3d_mlp_get(self.h, p.ctypes.data, C.byref(nrm)))
        return p, nrm

    def mlp_set(self, params, norm):
        import numpy as np
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_NPARAM
        if isinstance(norm, dict):
            norm = MlpNorm.from_numpy(norm)
        return check("c3d_mlp_set", self.L.c3d_mlp_set(self.h, p.ctypes.data, C.byref(norm)))

    def set_warmstart(self, on):
        return check("c3d_set_warmstart", self.L.c3d_set_warmstart(self.h, int(on)))

    def mlp_features(self, u, norm, X, alpha=None, T=None, stream=None):
        n = u.numel() // self.nm
        if isinstance(norm, dict):
            norm = MlpNorm.from_numpy(norm)
        return check("c3d_mlp_features", self.L.c3d_mlp_features(
            self.h, n, _p(u), _p(alpha), C.byref(norm), _p(X), _p(T), _stream(stream)))

    def mlp_forward(self, params, X, Y, stream=None):
        n = X.numel() // MLP_IN
        return check("c3d_mlp_forward", self.L.c3d_mlp_forward(self.h, n, _p(params), _p(X),
                                                               _p(Y), _stream(stream)))

    def mlp_infer(self, X, Yd, stream=None):
        n = X.numel() // MLP_IN
        return check("c3d_mlp_infer", self.L.c3d_mlp_infer(self.h, n, _p(X), _p(Yd),
                                                           _stream(stream)))

    def mlp_loss_grad(self, params, X, T, grad, loss, stream=None):
        n = X.numel() // MLP_IN
        return check("c3d_mlp_loss_grad", self.L.c3d_mlp_loss_grad(
            self.h, n, _p(params), _p(X), _p(T), _p(grad), _p(loss), _stream(stream)))


STATS_FIELDS = [("r", "f8"), ("err", "f8"), ("gamma", "f8"), ("totalIter", "i4"),
                ("iterGamma", "i4"), ("regIndex", "i4"), ("flags", "i4")]
STATS_BYTES = 40  # sizeof(c3d_cell_stats)


def alloc_stats(n, device):
    """A zeroed device buffer for n c3d_cell_stats records."""
    import torch
    return torch.zeros(n * STATS_BYTES, dtype=torch.uint8, device=device)


def stats_to_numpy(t):
    """A device buffer of ncells c3d_cell_stats records -> numpy structured array."""
    import numpy as np
    raw = t.detach().cpu().contiguous().view(-1).numpy().view(np.uint8)
    return raw.view(np.dtype(STATS_FIELDS))


def dense_solve(nm, H, g, d, rcond, info, L, lapack_order=0, stream=None):
    """Batched Cholesky + dpocon + Newton direction on device tensors."""
    n = H.numel() // (nm * nm)
    return check("c3d_dense_solve", load().c3d_dense_solve(
        nm, n, _p(H), _p(g), _p(d), _p(rcond), _p(info), _p(L), int(lapack_order),
        _stream(stream)))


def tables_host(N, Q):
    """The library's own tables (host only; no device needed)."""
    import numpy as np
    L = load()
    nm, nm2, nq = (N + 1) ** 2, (2 * N + 1) ** 2, 2 * Q * Q
    out = {k: np.zeros(nq) for k in ("w", "xi", "eta", "mu")}
    out["M"] = np.zeros((nq, nm))
    out["M2"] = np.zeros((nq, nm2))
    out["beta"] = np.zeros((nm * nm, nm2))
    out["pnOps"] = np.zeros((6, nm, nm))
    nnz = C.c_int()
    rowptr = np.zeros(nm * nm + 1, dtype=np.int32)
    check("c3d_tables_host", L.c3d_tables_host(
        N, Q, out["w"].ctypes.data, out["xi"].ctypes.data, out["eta"].ctypes.data,
        out["mu"].ctypes.data, out["M"].ctypes.data, out["M2"].ctypes.data,
        out["beta"].ctypes.data, out["pnOps"].ctypes.data, C.byref(nnz), 0, None, None,
        rowptr.ctypes.data))
    val = np.zeros(max(nnz.value, 1))
    col = np.zeros(max(nnz.value, 1), dtype=np.int32)
    check("c3d_tables_host", L.c3d_tables_host(
        N, Q, None, None, None, None, None, None, None, None, C.byref(nnz), nnz.value,
        val.ctypes.data, col.ctypes.data, None))
    out["nnz"] = nnz.value
    out["betaVal"], out["betaColInd"], out["betaRowPtr"] = val[:nnz.value], col[:nnz.value], rowptr
    return out
