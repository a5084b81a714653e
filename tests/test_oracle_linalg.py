"""Pins the oracle's LAPACK/BLAS restatements (dpotrf 'L', dpocon 'L', dpotrs, dnrm2; called
at optBatch.cpp:199-261 and opt.cpp:227-291), the custom solve of utils.cpp:71-86, norm1,
modcfl, and the deterministic exp/log, against scipy's LAPACK (live and committed fixtures),
numpy, and closed forms."""
import math
import os

import numpy as np
import pytest
import scipy.linalg.lapack as la

from conftest import ROOT

GOLD = np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))


def _spd(rng, n, logcond):
    A = rng.normal(size=(n, n))
    Q, _ = np.linalg.qr(A)
    H = (Q * np.logspace(0, -logcond, n)) @ Q.T
    return 0.5 * (H + H.T)


def _potrf(O, H):
    n = H.shape[0]
    A = np.ascontiguousarray(H.T).copy()  # column-major storage of H
    info = O.lib().orc_dpotrf_l(n, A)
    return A, info


def _colmajor_lower(A, n):
    return np.tril(A.reshape(n, n).T)


@pytest.mark.parametrize("n", [4, 9, 16, 25])
def test_dpotrf_against_scipy(O, n):
    rng = np.random.default_rng(n)
    for k in range(0, 14, 2):
        H = _spd(rng, n, k)
        A, info = _potrf(O, H)
        c, inf = la.dpotrf(H, lower=1)
        assert info == inf == 0
        L = _colmajor_lower(A, n)
        # forward error of a Cholesky factor grows like cond(H) * eps
        assert np.abs(L - np.tril(c)).max() <= 4e-16 * n * 10 ** k * np.abs(c).max()
        assert np.abs(L @ L.T - H).max() <= 1e-15 * np.abs(H).max() * n
    H = np.eye(n)
    H[2, 2] = -1
    _, info = _potrf(O, H)
    assert info == la.dpotrf(H, lower=1)[1] == 3


def test_dpocon_against_fixtures(O):
    L = O.lib()
    Hs, rc, info = GOLD["pocon_H"], GOLD["pocon_rcond"], GOLD["potrf_info"]
    for H25, (nm, r), (_, inf) in zip(Hs, rc, info):
        n = int(nm)
        H = np.ascontiguousarray(H25[:n, :n])
        A, i = _potrf(O, H)
        assert i == inf
        if i:
            continue
        ci = O.C.c_int()
        rcond = L.orc_dpocon_l(n, A, L.orc_norm1(n, H), O.C.byref(ci))
        # scipy's LAPACK (OpenBLAS) factors and solves in another order; the estimates
        # agree to ~cond(H) * eps relative
        assert abs(rcond - r) <= 1e-12 * r + 2e-15 * n, (n, rcond, r)


@pytest.mark.parametrize("n", [4, 9, 16, 25])
def test_dpocon_live_scipy(O, n):
    L = O.lib()
    rng = np.random.default_rng(100 + n)
    for trial in range(40):
        H = _spd(rng, n, rng.uniform(0, 13))
        A, i = _potrf(O, H)
        ci = O.C.c_int()
        anorm = np.abs(H).sum(axis=1).max()
        assert abs(L.orc_norm1(n, np.ascontiguousarray(H)) - anorm) <= 1e-15 * anorm
        r = L.orc_dpocon_l(n, A, anorm, O.C.byref(ci))
        rs, _ = la.dpocon(la.dpotrf(H, lower=1)[0], anorm, uplo="L")
        assert abs(r - rs) <= 1e-12 * rs + 2e-15 * n, (trial, r, rs)
        # the estimate bounds the true 1-norm condition number from below, within 3x
        true = 1.0 / (np.linalg.norm(H, 1) * np.linalg.norm(np.linalg.inv(H), 1))
        assert r >= true * 0.999 and r <= 3.0 * true * 10


@pytest.mark.parametrize("n", [4, 16, 25])
def test_dpotrs_variants(O, n):
    L = O.lib()
    rng = np.random.default_rng(7 * n)
    for k in (0, 4, 8):
        H = _spd(rng, n, k)
        A, _ = _potrf(O, H)
        g = rng.normal(size=n)
        ref = np.linalg.solve(H, g)
        for fn in (L.orc_dpotrs_custom, L.orc_dpotrs_lapack):
            d = g.copy()
            fn(n, A, d)
            assert np.abs(d - ref).max() <= 1e-15 * 10 ** k * np.abs(ref).max() * n
        c, _ = la.dpotrf(H, lower=1)
        x, _ = la.dpotrs(c, g, lower=1)
        d = g.copy()
        L.orc_dpotrs_lapack(n, A, d)
        assert np.abs(d - x).max() <= 1e-14 * 10 ** k * np.abs(x).max()


def test_dnrm2_norm1_modcfl(O):
    L = O.lib()
    rng = np.random.default_rng(3)
    for n in (1, 4, 9, 16, 25):
        for scale in (1e-200, 1e-8, 1.0, 1e150):
            x = rng.normal(size=n) * scale
            v = L.orc_dnrm2(n, x)
            ref = np.linalg.norm(x / scale) * scale  # numpy squares without scaling
            assert abs(v - ref) <= 4e-16 * ref
    assert L.orc_dnrm2(3, np.zeros(3)) == 0.0
    # modcfl: exp(2 ||d||_1 sqrt((2p+1)/(2pi))), p = floor(sqrt(n)) + 1
    for n in (4, 9, 16, 25):
        d = rng.normal(size=n) * 1e-3
        p = int(math.floor(math.sqrt(n))) + 1
        ref = math.exp(2.0 * np.abs(d).sum() * math.sqrt((2 * p + 1) / (2 * math.pi)))
        assert abs(L.orc_modcfl(n, d) - ref) <= 2e-16 * ref


def test_deterministic_exp_log(O):
    L = O.lib()
    rng = np.random.default_rng(11)
    xs = np.concatenate([rng.uniform(-708, 709, 20000), rng.uniform(-40, 40, 20000),
                         [0.0, -0.0, 1.0, -1.0, 1e-300, -1e-300]])
    e = np.array([L.orc_exp_det(float(x)) for x in xs])
    assert np.all(np.abs(e - np.exp(xs)) <= np.spacing(np.exp(xs)))  # within 1 ulp
    assert L.orc_exp_det(0.0) == 1.0
    assert L.orc_exp_det(800.0) == math.inf and L.orc_exp_det(-800.0) == 0.0
    assert math.isnan(L.orc_exp_det(math.nan))
    ys = np.exp(rng.uniform(-700, 700, 20000))
    lg = np.array([L.orc_log_det(float(y)) for y in ys])
    assert np.all(np.abs(lg - np.log(ys)) <= 2 * np.spacing(np.abs(np.log(ys))) + 1e-300)
    assert L.orc_log_det(1.0) == 0.0 and L.orc_log_det(0.0) == -math.inf
