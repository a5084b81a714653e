"""Pins the oracle's quadrature / basis / Gaunt / P_N tables (init.cpp:423-548, 287-337,
199-234; utils.cpp:52-68) and checks the product's own table builder against them.

Pins: committed numpy/scipy fixtures (tests/golden/golden.npz), closed forms for n <= 1,
orthonormality under the quadrature, numpy least squares and the exactness identity for
the Gaunt coefficients."""
import math
import os

import numpy as np
import pytest

from conftest import ROOT, product

GOLD = np.load(os.path.join(ROOT, "tests", "golden", "golden.npz"))


@pytest.mark.parametrize("Q", [1, 2, 3, 5, 8, 12, 24, 40])
def test_gauss_legendre_matches_fixture(O, Q):
    x = np.zeros(Q)
    w = np.zeros(Q)
    O.lib().orc_gauss_legendre(Q, x, w)
    assert np.all(np.diff(x) > 0), "ascending, as GSL's glfixed orders them"
    assert np.allclose(x, GOLD[f"gl_Q{Q}_x"], rtol=0, atol=2e-16 * 4)
    # numpy's leggauss and scipy's roots_legendre weights disagree with each other by up to
    # 3e-13 relative at Q = 40; the polynomial exactness loop below is the sharper pin
    assert np.allclose(w, GOLD[f"gl_Q{Q}_w"], rtol=1e-12, atol=0)
    assert abs(w.sum() - 2.0) < 1e-14
    # exact for polynomials of degree 2Q-1
    for deg in range(2 * Q):
        exact = 0.0 if deg % 2 else 2.0 / (deg + 1)
        assert abs(np.sum(w * x ** deg) - exact) < 1e-13


def test_sph_plm_matches_fixture(O):
    L = O.lib()
    for l, m, x, ref in GOLD["plm"]:
        v = L.orc_sph_plm(int(l), int(m), float(x))
        assert abs(v - ref) <= 1e-13 * max(1.0, abs(ref)), (l, m, x, v, ref)


def test_basis_closed_forms_and_orthonormality(O):
    T = O.Tables(4, 10)
    m0 = 1.0 / (2.0 * math.sqrt(math.pi))
    assert np.allclose(T.M[:, 0], m0, rtol=1e-15, atol=0)
    c = math.sqrt(3.0 / (4.0 * math.pi))
    assert np.allclose(T.M[:, 1], c * T.eta, atol=1e-15)
    assert np.allclose(T.M[:, 2], c * T.mu, atol=1e-15)
    assert np.allclose(T.M[:, 3], -c * T.xi, atol=1e-15)
    G = (T.M * T.w[:, None]).T @ T.M
    assert np.abs(G - np.eye(T.nm)).max() < 1e-14
    G2 = (T.M2 * T.w[:, None]).T @ T.M2
    assert np.abs(G2 - np.eye(T.nm2)).max() < 1e-13
    assert abs(T.w.sum() - 4 * math.pi) < 1e-13
    assert np.allclose(T.xi ** 2 + T.eta ** 2 + T.mu ** 2, 1.0, atol=1e-15)
    # the first nm functions of the extended basis are the basis itself (bitwise)
    assert np.array_equal(T.M2[:, :T.nm], T.M)


@pytest.mark.parametrize("N,Q", [(1, 4), (2, 8), (3, 12)])
def test_gaunt_least_squares_and_exactness(O, N, Q):
    T = O.Tables(N, Q)
    for ij in range(0, T.nm * T.nm, 7):
        i, j = divmod(ij, T.nm)
        b = T.M2[:, i] * T.M2[:, j]
        ref = np.linalg.lstsq(T.M2, b, rcond=None)[0]
        assert np.abs(T.beta[ij] - ref).max() < 1e-13
    recon = np.einsum("pk,qk->pq", T.beta, T.M2).reshape(T.nm, T.nm, T.nq)
    prod = T.M[:, :, None] * T.M[:, None, :]
    assert np.abs(recon - np.transpose(prod, (1, 2, 0))).max() < 1e-13
    # CSR keeps exactly the entries above 1e-13
    assert T.betaRowPtr[-1] == T.nnz == int((np.abs(T.beta) > 1e-13).sum())


def test_pn_operators(O):
    T = O.Tables(2, 8)
    full = {}
    for o, dirv in enumerate((T.xi, T.eta, T.mu)):
        plus = dirv > 0
        for half, sel in ((0, plus), (1, ~plus)):
            ref = np.einsum("q,qi,qj->ij", (dirv * T.w)[sel], T.M[sel], T.M[sel])
            assert np.abs(T.ops[2 * o + half] - ref).max() < 1e-14
        full[o] = T.ops[2 * o] + T.ops[2 * o + 1]
        assert np.abs(full[o] - full[o].T).max() < 1e-15


@pytest.mark.parametrize("N,Q", [(1, 6), (2, 8), (3, 24), (4, 40)])
def test_product_tables_match_oracle(O, N, Q):
    P = product()
    t = P.tables_host(N, Q)
    T = O.Tables(N, Q)
    for k in ("w", "xi", "eta", "mu", "M", "M2"):
        ref = getattr(T, k)
        assert np.abs(t[k] - ref).max() <= 16e-16 * max(1.0, np.abs(ref).max()), k
    assert np.abs(t["beta"] - T.beta).max() < 1e-13
    assert t["nnz"] == T.nnz
    assert np.array_equal(t["betaRowPtr"], T.betaRowPtr)
    assert np.array_equal(t["betaColInd"], T.betaColInd[:T.nnz])
    assert np.abs(t["pnOps"] - T.ops).max() < 1e-14


def test_gaunt_underdetermined_quadrature_is_declined(O):
    """Fewer ordinates than degree-2N basis functions (N = 4, Q = 6: 72 < 81): the least
    squares of init.cpp:287-337 is underdetermined; the oracle and the library both build no
    coefficients (nnz = 0, beta zero) instead of writing past their work arrays."""
    import numpy as np
    from conftest import product
    T = O.Tables(4, 6)
    assert T.nq < T.nm2 and T.nnz == 0 and not T.beta.any()
    h = product().tables_host(4, 6)
    assert h["nnz"] == 0 and not h["beta"].any()
    assert np.allclose(h["w"].sum(), 4 * np.pi)
