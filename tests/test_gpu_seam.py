"""GPU parity of the accelerator seam, the drop-ins for src/cuda/cuda_headers.h:10-21, called
as the reference calls them -- host arrays, synchronous -- through ctypes:

  solveHInit_hip / solveH_hip   (optBatch.cpp:364-393): a batch of Hessians from the caller's
                                tables, direct and Gaunt form, against the oracle's solveHMn;
                                repeated calls of one size reuse the staging buffers.
  solveFluxKinetic_hip          (flux_kinetic.cpp:84-92): grid sizes including the ghosts,
                                theta = 2 (cuda_flux_kinetic.cu:82), M_N and P_N ansatz, on a
                                grid with nx != nz (the reference kernel's defect,
                                cuda_flux_kinetic.cu:150-151, is not reproduced), against the
                                oracle's solveFluxKinetic; ghost cells of the flux are zero."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("N,Q,gaunt", [(2, 8, 0), (3, 12, 0), (2, 8, 1), (3, 12, 1)])
def test_solveH_seam_matches_oracle(P, O, dev, N, Q, gaunt):
    L = P.load()
    T = O.Tables(N, Q)
    nm, nm2, nq = T.nm, T.nm2, T.nq
    batch = 50  # the reference deck's batch size (input.deck:54)
    rng = np.random.default_rng(N + 10 * gaunt)
    alphas = rng.normal(0, 0.4, (batch, nm))
    alphas[:, 0] = rng.uniform(-2.0, 1.0, batch)
    L.solveHInit_hip(nm, nm2, nq, batch, _p(T.w), _p(T.M), _p(T.M2), _p(T.beta),
                     _p(T.betaVal), _p(T.betaColInd), _p(T.betaRowPtr), gaunt)
    assert L.c3d_seam_status() == 0
    ref = np.zeros((batch, nm, nm))
    for b in range(batch):
        O.lib().orc_solveH(nm, nm2, nq, alphas[b], O.C.byref(T.c), gaunt, 0, ref[b],
                           np.zeros(nm2))
    for call in range(3):  # the staging buffers are reused
        hs = np.full((batch, nm, nm), np.nan)
        L.solveH_hip(nm, nm2, nq, batch, _p(alphas), _p(hs), gaunt, 0)
        assert L.c3d_seam_status() == 0
        assert np.isfinite(hs).all()
        assert np.array_equal(hs, np.transpose(hs, (0, 2, 1)))  # mirrored, as solveHMn
        assert np.abs(hs - ref).max() <= 1e-13 * np.abs(ref).max()
    # a smaller batch after larger ones
    hs = np.zeros((7, nm, nm))
    L.solveH_hip(nm, nm2, nq, 7, _p(alphas[:7].copy()), _p(hs), gaunt, 0)
    assert L.c3d_seam_status() == 0
    assert np.abs(hs - ref[:7]).max() <= 1e-13 * np.abs(ref).max()


def test_solveH_seam_reports_uninitialised_sizes(P, O, dev):
    L = P.load()
    T = O.Tables(1, 6)
    L.solveHInit_hip(T.nm, T.nm2, T.nq, 4, _p(T.w), _p(T.M), _p(T.M2), None, None, None,
                     None, 0)
    assert L.c3d_seam_status() == 0
    hs = np.zeros((4, 9, 9))
    a = np.zeros((4, 9))
    L.solveH_hip(9, 25, T.nq, 4, _p(a), _p(hs), 0, 0)  # nm of another order
    assert L.c3d_seam_status() == P.C3D_ERR_ARG


@pytest.mark.parametrize("isPn,dims", [(0, (9, 8, 6)), (1, (9, 8, 6)), (0, (6, 7, 10))])
def test_solveFluxKinetic_seam_matches_oracle(P, O, dev, isPn, dims):
    L = P.load()
    T = O.Tables(2, 8)
    nx, ny, nz = dims
    u, _, _ = O.init_grid(nx, ny, nz, T.nm, ic=1, sigma=0.3, floor=1e-3, noise=0.05, seed=4)
    rng = np.random.default_rng(4)
    if isPn:
        alpha = u.copy()
        alpha[:, 1:] = u[:, :1] * rng.uniform(-0.2, 0.2, (u.shape[0], T.nm - 1))
    else:
        alpha = O.init_alpha(u)
        alpha[:, 1:] = rng.normal(0, 0.3, (u.shape[0], T.nm - 1))
    d = (2 / nx, 2 / ny, 2 / nz)
    ng = u.shape[0]
    ref = np.zeros_like(u)
    O.lib().orc_flux_kinetic(nx, ny, nz, T.nm, T.nq, alpha, ref, O.C.byref(T.c), d[0], d[1],
                             d[2], 2.0, isPn, np.zeros(ng * T.nq), 0)
    for call in range(2):
        flux = np.full_like(u, np.nan)
        L.solveFluxKinetic_hip(nx + 4, ny + 4, nz + 4, T.nm, T.nq, d[0], d[1], d[2], _p(alpha),
                               _p(flux), _p(T.xi), _p(T.eta), _p(T.mu), _p(T.w), _p(T.M), isPn)
        assert L.c3d_seam_status() == 0
        F = flux.reshape(nx + 4, ny + 4, nz + 4, T.nm)
        R = ref.reshape(nx + 4, ny + 4, nz + 4, T.nm)
        inner = (slice(2, -2),) * 3
        if isPn:
            assert np.array_equal(F[inner], R[inner])
        else:
            assert np.abs(F[inner] - R[inner]).max() <= 1e-12 * np.abs(R[inner]).max()
        ghost = np.ones(F.shape[:3], bool)
        ghost[inner] = False
        assert (F[ghost] == 0).all()
