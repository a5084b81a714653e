"""Pins the oracle bit for bit against the reference's own CPU sources, compiled where they
lie into oracle/_ref/libref.so (src/opt/fobj.cpp, src/opt/linesearch.cpp,
src/flux_kinetic.cpp, src/flux_moment.cpp, src/comm.cpp serial branch, src/global.cpp).
Both sides use the C library's exp (oracle exp mode 0) and are built without FMA
contraction, so the results must be identical, not merely close."""
import numpy as np
import pytest

import oracle as O_mod

pytestmark = pytest.mark.skipif(not O_mod.Ref.available,
                                reason="oracle/_ref/libref.so not built (reference absent)")


@pytest.mark.parametrize("N,Q", [(1, 6), (2, 8), (3, 12), (4, 10)])
def test_solveH_and_fobj_bitwise(O, N, Q):
    T = O.Tables(N, Q)
    R = O.Ref(T)
    rng = np.random.default_rng(N)
    gx = np.zeros(T.nm2)
    for trial in range(6):
        a = rng.normal(0, 0.8, T.nm)
        u = rng.normal(0, 0.2, T.nm)
        u[0] = 1.0
        for gaunt, sparse in ((0, 0), (1, 0), (1, 1)):
            h1 = np.zeros((T.nm, T.nm))
            h2 = np.zeros((T.nm, T.nm))
            R.L.ref_solveHMn(a, h1, gaunt, sparse)
            O.lib().orc_solveH(T.nm, T.nm2, T.nq, a, O.C.byref(T.c), gaunt, sparse, h2, gx)
            assert np.array_equal(h1, h2), (gaunt, sparse)
        g1 = np.zeros(T.nm)
        g2 = np.zeros(T.nm)
        f1 = R.L.ref_fobjMn(a, u, g1.ctypes.data, None, 0, 0)
        f2 = O.lib().orc_fobj(T.nm, T.nq, a, u, T.w, T.M, g2.ctypes.data)
        assert f1 == f2 and np.array_equal(g1, g2)
        R.L.ref_solveHMn(a, h1, 0, 0)
        f1 = R.L.ref_fobjMnUseH(a, u, g1.ctypes.data, h1)
        f2 = O.lib().orc_fobj_useH(T.nm, a, u, T.M[0, 0], g2.ctypes.data, h1)
        assert f1 == f2 and np.array_equal(g1, g2)


def test_linesearch_bitwise(O):
    T = O.Tables(3, 12)
    R = O.Ref(T)
    rng = np.random.default_rng(5)
    accepted = 0
    for trial in range(40):
        a = rng.normal(0, 0.5, T.nm)
        u = rng.normal(0, 0.15, T.nm)
        u[0] = 1.0
        g = np.zeros(T.nm)
        f = O.lib().orc_fobj(T.nm, T.nq, a, u, T.w, T.M, g.ctypes.data)
        d = -g * rng.uniform(0.5, 30.0)  # long steps force backtracking
        a1, a2 = a.copy(), a.copy()
        t1 = R.L.ref_linesearch(a1, d, f, g, u)
        t2 = O.lib().orc_linesearch(T.nm, T.nq, a2, d, f, g, u, T.w, T.M)
        assert t1 == t2 and np.array_equal(a1, a2)
        accepted += t1 < 1.0
    assert accepted > 5


@pytest.mark.parametrize("isPn,theta", [(0, 2.0), (0, 1.0), (1, 2.0), (0, 1.5)])
def test_flux_kinetic_bitwise(O, isPn, theta):
    T = O.Tables(2, 8)
    nx, ny, nz = 7, 6, 5
    R = O.Ref(T, nx, ny, nz, theta)
    u, _, _ = O.init_grid(nx, ny, nz, T.nm, ic=1, sigma=0.3, floor=1e-3, noise=0.05, seed=2)
    rng = np.random.default_rng(4)
    u[:, 1:] = u[:, :1] * rng.uniform(-0.2, 0.2, (u.shape[0], T.nm - 1))
    alpha = u.copy() if isPn else O.init_alpha(u) + np.pad(
        rng.normal(0, 0.3, (u.shape[0], T.nm - 1)), ((0, 0), (1, 0)))
    f1 = np.full_like(u, 7.0)
    f2 = np.full_like(u, 7.0)
    R.L.ref_solveFluxKinetic(u, f1, alpha, 2 / nx, 2 / ny, 2 / nz, isPn)
    O.lib().orc_flux_kinetic(nx, ny, nz, T.nm, T.nq, alpha, f2, O.C.byref(T.c), 2 / nx, 2 / ny,
                             2 / nz, theta, isPn, np.zeros(u.shape[0]), 0)
    assert np.array_equal(f1, f2)


def test_flux_moment_and_ghost_fill_bitwise(O):
    T = O.Tables(2, 8)
    nx, ny, nz = 6, 7, 8
    R = O.Ref(T, nx, ny, nz)
    u, _, _ = O.init_grid(nx, ny, nz, T.nm, ic=1, sigma=0.3, floor=1e-3, noise=0.05, seed=9)
    rng = np.random.default_rng(6)
    u[:, 1:] = u[:, :1] * rng.uniform(-0.2, 0.2, (u.shape[0], T.nm - 1))
    f1 = np.full_like(u, 3.0)
    f2 = np.full_like(u, 3.0)
    R.L.ref_solveFluxMoment(u, f1, 0.2, 0.25, 0.3)
    O.lib().orc_flux_moment(nx, ny, nz, T.nm, u, f2, T.ops, 0.2, 0.25, 0.3, 0)
    assert np.array_equal(f1, f2)
    u1, u2 = u.copy(), u.copy()
    R.L.ref_communicateBoundaries(u1)
    O.lib().orc_ghost_fill(nx, ny, nz, T.nm, u2)
    assert np.array_equal(u1, u2) and not np.array_equal(u1, u)
