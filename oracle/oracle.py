"""ctypes front end of the CPU oracle (oracle/liboracle.so) and of the partial reference
build (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline.  The product never imports it.
"""
import ctypes as C
import os
import shutil
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_vp = C.c_void_p

FLAG_ISO_RESTART = 1
FLAG_U0_NONPOS = 2
FLAG_LADDER = 4
REGULARIZATIONS = (0.0, 1e-3, 1e-2, 1e-1, 5e-1)

STATS_DTYPE = np.dtype([("r", "f8"), ("err", "f8"), ("gamma", "f8"), ("totalIter", "i4"),
                        ("iterGamma", "i4"), ("regIndex", "i4"), ("flags", "i4")])


class Opts(C.Structure):
    _fields_ = [("maxIter", C.c_int32), ("useGaunt", C.c_int32),
                ("useGauntSparse", C.c_int32), ("unbatched", C.c_int32),
                ("tol", C.c_double), ("tolGamma", C.c_double), ("condHMax", C.c_double)]


class TablesC(C.Structure):
    _fields_ = [("N", C.c_int32), ("Q", C.c_int32), ("nm", C.c_int32), ("nm2", C.c_int32),
                ("nq", C.c_int32), ("nnz", C.c_int32),
                ("w", _vp), ("xi", _vp), ("eta", _vp), ("mu", _vp), ("M", _vp), ("M2", _vp),
                ("beta", _vp), ("betaVal", _vp), ("betaColInd", _vp), ("betaRowPtr", _vp)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or build())")
        L = C.CDLL(LIB_PATH)
        L.orc_set_exp_mode.argtypes = [C.c_int]
        L.orc_exp_det.argtypes = [C.c_double]
        L.orc_exp_det.restype = C.c_double
        L.orc_log_det.argtypes = [C.c_double]
        L.orc_log_det.restype = C.c_double
        L.orc_gauss_legendre.argtypes = [C.c_int, _dp, _dp]
        L.orc_sph_plm.argtypes = [C.c_int, C.c_int, C.c_double]
        L.orc_sph_plm.restype = C.c_double
        L.orc_real_sh.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_real_sh.restype = C.c_double
        L.orc_quadrature.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_gaunt.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _vp, _vp, _vp]
        L.orc_gaunt.restype = C.c_int
        L.orc_pn_operators.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_dpotrf_l.argtypes = [C.c_int, _dp]
        L.orc_dpotrf_l.restype = C.c_int
        L.orc_dpocon_l.argtypes = [C.c_int, _dp, C.c_double, C.POINTER(C.c_int)]
        L.orc_dpocon_l.restype = C.c_double
        L.orc_dpotrs_custom.argtypes = [C.c_int, _dp, _dp]
        L.orc_dpotrs_lapack.argtypes = [C.c_int, _dp, _dp]
        L.orc_dnrm2.argtypes = [C.c_int, _dp]
        L.orc_dnrm2.restype = C.c_double
        L.orc_norm1.argtypes = [C.c_int, _dp]
        L.orc_norm1.restype = C.c_double
        L.orc_modcfl.argtypes = [C.c_int, _dp]
        L.orc_modcfl.restype = C.c_double
        L.orc_solveH.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.POINTER(TablesC), C.c_int,
                                 C.c_int, _dp, _dp]
        L.orc_fobj.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _vp]
        L.orc_fobj.restype = C.c_double
        L.orc_fobj_useH.argtypes = [C.c_int, _dp, _dp, C.c_double, _vp, _dp]
        L.orc_fobj_useH.restype = C.c_double
        L.orc_linesearch.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_linesearch.restype = C.c_double
        L.orc_closure_cell.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.POINTER(Opts),
                                       C.POINTER(TablesC), _vp]
        L.orc_closure_cell.restype = C.c_int
        L.orc_closure.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                  C.POINTER(Opts), C.POINTER(TablesC), _vp, C.c_int]
        L.orc_closure.restype = C.c_int
        L.orc_ghost_fill.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_flux_kinetic.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp,
                                       C.POINTER(TablesC), C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_int, _dp, C.c_int]
        L.orc_flux_moment.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                      C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_euler.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp,
                                C.c_double]
        L.orc_update.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                 C.POINTER(TablesC), _dp, C.POINTER(Opts), C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.c_double, _vp, C.c_int]
        L.orc_update.restype = C.c_int
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_uniform.restype = C.c_double
        L.orc_normal.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_normal.restype = C.c_double
        L.orc_init_grid.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                    C.c_double, C.c_double, C.c_double, C.c_uint64,
                                    C.c_double, _dp, _dp, _dp]
        L.orc_init_alpha.argtypes = [C.c_int64, C.c_int, _dp, _dp]
        L.orc_closure_batch_input.argtypes = [C.c_int64, C.c_int, C.c_int, _dp, _dp, C.c_uint64,
                                              C.c_double, C.c_double, _dp, _dp, _dp]
        _lib = L
    return _lib


class exp_mode:
    """Context manager selecting the oracle's exp/log: 0 = C library, 1 = deterministic."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        lib().orc_set_exp_mode(self.mode)
        return self

    def __exit__(self, *a):
        lib().orc_set_exp_mode(0)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Tables:
    """Quadrature / basis / Gaunt tables of init.cpp:423-548 and :287-337."""

    def __init__(self, N, Q, gaunt=True):
        L = lib()
        self.N, self.Q = N, Q
        self.nm, self.nm2, self.nq = (N + 1) ** 2, (2 * N + 1) ** 2, 2 * Q * Q
        nq, nm, nm2 = self.nq, self.nm, self.nm2
        self.w, self.xi, self.eta, self.mu = (np.zeros(nq) for _ in range(4))
        self.M = np.zeros((nq, nm))
        self.M2 = np.zeros((nq, nm2))
        L.orc_quadrature(N, Q, self.w, self.xi, self.eta, self.mu, self.M, self.M2)
        self.beta = np.zeros((nm * nm, nm2))
        if gaunt:
            nnz = L.orc_gaunt(nm, nm2, nq, self.M2, self.beta, None, None, None)
            self.betaVal = np.zeros(max(nnz, 1))
            self.betaColInd = np.zeros(max(nnz, 1), dtype=np.int32)
            self.betaRowPtr = np.zeros(nm * nm + 1, dtype=np.int32)
            L.orc_gaunt(nm, nm2, nq, self.M2, self.beta, _ptr(self.betaVal),
                        _ptr(self.betaColInd), _ptr(self.betaRowPtr))
            self.nnz = nnz
        else:
            self.betaVal = np.zeros(1)
            self.betaColInd = np.zeros(1, dtype=np.int32)
            self.betaRowPtr = np.zeros(nm * nm + 1, dtype=np.int32)
            self.nnz = 0
        self.ops = np.zeros((6, nm, nm))
        L.orc_pn_operators(nm, nq, self.w, self.xi, self.eta, self.mu, self.M, self.ops)
        self.c = TablesC(N, Q, nm, nm2, nq, self.nnz, _ptr(self.w), _ptr(self.xi),
                         _ptr(self.eta), _ptr(self.mu), _ptr(self.M), _ptr(self.M2),
                         _ptr(self.beta), _ptr(self.betaVal), _ptr(self.betaColInd),
                         _ptr(self.betaRowPtr))


def opts(maxIter=100, tol=1e-8, tolGamma=1.1, condHMax=1e10, useGaunt=False,
         useGauntSparse=False, unbatched=False):
    return Opts(maxIter, int(useGaunt), int(useGauntSparse), int(unbatched), tol, tolGamma,
                condHMax)


def closure(T, u, alpha, o=None, nthreads=0):
    """Solve the closure in place on AoS arrays u, alpha of shape (ncells, nm).
    Returns (status_bits, stats structured array)."""
    if o is None:
        o = opts()
    n = u.shape[0]
    stats = np.zeros(n, dtype=STATS_DTYPE)
    rc = lib().orc_closure(n, T.nm, T.nm2, T.nq, u, alpha, C.byref(o), C.byref(T.c),
                           stats.ctypes.data_as(C.c_void_p), nthreads)
    return rc, stats


def closure_batch_input(T, n, seed, smin=0.5, smax=8.0):
    a_star = np.zeros((n, T.nm))
    u = np.zeros((n, T.nm))
    a0 = np.zeros((n, T.nm))
    lib().orc_closure_batch_input(n, T.nm, T.nq, T.w, T.M, seed, smin, smax, a_star, u, a0)
    return a_star, u, a0


def grid_shape(nx, ny, nz):
    return (nx + 4, ny + 4, nz + 4)


def init_grid(nx, ny, nz, nm, a=(-1.0, -1.0, -1.0), d=None, ic=1, sigma=0.1, floor=1e-4,
              noise=0.0, seed=0, cross=1.0):
    if d is None:
        d = ((1.0 - a[0]) / nx, (1.0 - a[1]) / ny, (1.0 - a[2]) / nz)
    ng = (nx + 4) * (ny + 4) * (nz + 4)
    u = np.zeros((ng, nm))
    sT = np.zeros(ng)
    sS = np.zeros(ng)
    lib().orc_init_grid(nx, ny, nz, nm, a[0], a[1], a[2], d[0], d[1], d[2], ic, sigma, floor,
                        noise, seed, cross, u, sT, sS)
    return u, sT, sS


def init_alpha(u):
    a = np.zeros_like(u)
    lib().orc_init_alpha(u.shape[0], u.shape[1], u, a)
    return a


class Ref:
    """A fresh instance of the partial reference build (oracle/_ref/libref.so).

    Each instance loads its own copy of the library, because the reference keeps
    first-call static buffers sized by the first grid it sees."""

    available = os.path.exists(REF_PATH)

    def __init__(self, T, nx=1, ny=1, nz=1, theta=2.0):
        if not os.path.exists(REF_PATH):
            raise RuntimeError("reference build missing (oracle/_ref/libref.so)")
        self._dir = tempfile.mkdtemp(prefix="c3dref_")
        path = os.path.join(self._dir, "libref.so")
        shutil.copyfile(REF_PATH, path)
        L = C.CDLL(path)
        self.L = L
        self.T = T
        L.ref_setup.argtypes = [C.c_int] * 6 + [C.c_double] + [_vp] * 11
        L.ref_solveHMn.argtypes = [_dp, _dp, C.c_int, C.c_int]
        L.ref_fobjMn.argtypes = [_dp, _dp, _vp, _vp, C.c_int, C.c_int]
        L.ref_fobjMn.restype = C.c_double
        L.ref_fobjMnUseH.argtypes = [_dp, _dp, _vp, _dp]
        L.ref_fobjMnUseH.restype = C.c_double
        L.ref_linesearch.argtypes = [_dp, _dp, C.c_double, _dp, _dp]
        L.ref_linesearch.restype = C.c_double
        L.ref_solveFluxKinetic.argtypes = [_dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                           C.c_int]
        L.ref_solveFluxMoment.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double]
        L.ref_communicateBoundaries.argtypes = [_dp]
        L.ref_setup(nx, ny, nz, T.nm, T.nm2, T.nq, theta, _ptr(T.w), _ptr(T.xi), _ptr(T.eta),
                    _ptr(T.mu), _ptr(T.M), _ptr(T.M2), _ptr(T.beta), _ptr(T.betaVal),
                    _ptr(T.betaColInd), _ptr(T.betaRowPtr), _ptr(T.ops))

    def __del__(self):
        try:
            shutil.rmtree(self._dir, ignore_errors=True)
        except Exception:
            pass
