"""Pins of the oracle's stage update (orc_update / orc_euler, update.cpp:21-37, 185-275) and of
its unbatched closure branch (orc_closure_cell with unbatched = 1, opt.cpp:139-355) by facts
the mathematics and the physics fix, so that a dropped term, a flipped sign, a wrong stage
or a wrong stopping rule fails one of them:

  stage:    for P_N the semi-discrete system du/dt = L u is linear, so one SSP-RK2 (Heun)
            step is exactly u + dt L u + dt^2/2 L^2 u, with L = -(moment-flux divergence) +
            collisions (the flux is the reference's own solveFluxMoment, pinned bit for bit
            in test_oracle_ref.py); an eigenmode of A_x travels at its eigenvalue, in the
            direction of its sign; for M_N the density's centroid moves with the summed
            x-current (the discrete first-moment identity), and the step equals the
            reference's call sequence over the separately pinned closure, kinetic flux and
            Euler stage.
  closure:  unbatched cells stop on ||g||_2 <= ||u/u0||_2 tol (relative, recomputed here by
            numpy), some cells stop above the absolute tol, the last rung gets ten times
            maxIter, and a replay of optScaled built from the reference's compiled fobjMn
            and linesearch plus scipy's LAPACK reproduces the oracle's outcome tuples."""
import math

import numpy as np
import pytest
import scipy.linalg.lapack as la

import oracle as O_mod


def _interior(u, nx, ny, nz):
    return u.reshape(nx + 4, ny + 4, nz + 4, -1)[2:-2, 2:-2, 2:-2]


def _update(O, T, nx, ny, nz, u, alpha, sT, sS, d, dt, solver, pnk=0, theta=2.0, o=None):
    ng = u.shape[0]
    return O.lib().orc_update(nx, ny, nz, T.nm, T.nm2, T.nq, solver, pnk, u, alpha,
                              np.zeros_like(u), np.zeros_like(u), np.zeros(ng * T.nq), sT, sS,
                              O.C.byref(T.c), T.ops, O.C.byref(o or O.opts()), d[0], d[1], d[2],
                              theta, dt, None, 0)


# ------------------------------------------------------------------------- stage (P_N)

@pytest.mark.parametrize("N,dims,d", [(1, (5, 6, 4), (0.3, 0.25, 0.2)),
                                      (2, (6, 5, 7), (0.5, 0.125, 0.25))])
def test_pn_step_is_heun_of_the_linear_operator(O, N, dims, d):
    """orc_update (P_N, moment flux) against the closed form of Heun's method on the linear
    operator L v = -D(v) + C v: D the flux divergence of the periodic extension of v
    (solveFluxMoment), C the collision term of isotropic scattering (u0 gains sigma_s u0 and
    loses sigma_t u0, higher moments lose sigma_t u_m).  sigma_s != sigma_t per cell, so the
    two cross sections cannot be swapped unnoticed."""
    T = O.Tables(N, 8)
    L = O.lib()
    nx, ny, nz = dims
    ng = (nx + 4) * (ny + 4) * (nz + 4)
    rng = np.random.default_rng(11 + N)
    u = rng.normal(0.0, 1.0, (ng, T.nm))
    u[:, 0] += 3.0
    sT = rng.uniform(0.5, 1.5, ng)
    sS = rng.uniform(0.0, 1.0, ng)
    dt = 0.04

    def Lop(v):
        w = v.copy()
        L.orc_ghost_fill(nx, ny, nz, T.nm, w)
        F = np.zeros_like(w)
        L.orc_flux_moment(nx, ny, nz, T.nm, w, F, T.ops, d[0], d[1], d[2], 0)
        out = -F
        out[:, 0] += (sS - sT) * w[:, 0]
        out[:, 1:] -= sT[:, None] * w[:, 1:]
        r = np.zeros_like(out)  # the state space is the interior; ghosts follow from it
        _interior(r, nx, ny, nz)[...] = _interior(out, nx, ny, nz)
        return r

    u0 = np.zeros_like(u)
    _interior(u0, nx, ny, nz)[...] = _interior(u, nx, ny, nz)
    Lu = Lop(u0)
    expect = u0 + dt * Lu + 0.5 * dt * dt * Lop(Lu)
    got = u.copy()
    assert _update(O, T, nx, ny, nz, got, np.zeros_like(u), sT, sS, d, dt, solver=0) == 0
    gi, ei = _interior(got, nx, ny, nz), _interior(expect, nx, ny, nz)
    scale = np.abs(ei).max()
    assert np.abs(gi - ei).max() <= 1e-13 * scale
    # the closed form is sensitive to every piece: the step moves the state by O(dt)
    assert np.abs(gi - _interior(u0, nx, ny, nz)).max() > 1e-3 * scale


@pytest.mark.parametrize("which", ["max", "min"])
def test_pn_eigenmode_travels_at_its_eigenvalue(O, which):
    """u = v sin(k x) with A_x v = lambda v solves u_t + A_x u_x = 0 as v sin(k(x - lambda t)):
    the pattern must move by lambda t (sign included) -- a flipped flux sign moves it the
    other way, a wrong dt scale by the wrong distance."""
    T = O.Tables(1, 8)
    Ax = T.ops[0] + T.ops[1]  # A+ + A- = <xi m m^T>
    lam, V = np.linalg.eigh(Ax)
    kk = int(np.argmax(lam) if which == "max" else np.argmin(lam))
    v, lv = V[:, kk], lam[kk]
    assert abs(lv) > 0.5
    nx, ny, nz = 64, 4, 4
    d = (2.0 / nx, 0.5, 0.5)
    x = -1.0 + (np.arange(nx) + 0.5) * d[0]
    kw = math.pi
    ng = (nx + 4) * (ny + 4) * (nz + 4)
    u = np.zeros((ng, T.nm))
    U = _interior(u, nx, ny, nz)
    U[...] = np.sin(kw * x)[:, None, None, None] * v[None, None, None, :]
    zero = np.zeros(ng)
    dt = 0.4 * d[0]
    steps = 12
    for _ in range(steps):
        assert _update(O, T, nx, ny, nz, u, np.zeros_like(u), zero, zero, d, dt, solver=0) == 0
    a = (_interior(u, nx, ny, nz) @ v).mean(axis=(1, 2))
    c1, c2 = np.linalg.lstsq(np.c_[np.sin(kw * x), np.cos(kw * x)], a, rcond=None)[0]
    shift = math.atan2(-c2, c1) / kw
    want = lv * dt * steps
    assert abs(shift - want) <= 0.03 * abs(want), (shift, want)
    assert 0.95 <= math.hypot(c1, c2) <= 1.0 + 1e-12


# ------------------------------------------------------------------------- stage (M_N)

def _drifting_pulse(O, T, nx, ny, nz, beta0):
    """A realizable M_N state: density pulse in x on an isotropic floor, forward-peaked in +x
    inside the pulse (alpha_3 = -beta, m_3 = -sqrt(3/4pi) xi), uniform in y and z."""
    ng = (nx + 4) * (ny + 4) * (nz + 4)
    d = (2.0 / nx, 0.5, 0.5)
    x = -1.0 + (np.arange(nx + 4) - 2 + 0.5) * d[0]
    bump = np.exp(-x * x / (2 * 0.2 ** 2))
    alpha = np.zeros((nx + 4, ny + 4, nz + 4, T.nm))
    alpha[..., 0] = np.log(0.05 + bump)[:, None, None]
    alpha[..., 3] = (-beta0 * bump)[:, None, None]
    alpha = alpha.reshape(ng, T.nm)
    e = np.exp(alpha @ T.M.T) * T.w
    u = e @ T.M
    return u, alpha, d


def test_mn_centroid_moves_with_the_current(O):
    """First-moment identity: with sigma = 0 and zero current at the periodic boundary,
    d/dt sum_i x_i rho_i = sum_i J_x,i (rho = u0/m0, J_x = <xi f> = -u3 / sqrt(3/4pi)).
    One SSP-RK2 step of the oracle must move the centroid by dt times the mean current
    (to the scheme's O(dx^2)), in the direction of the current."""
    T = O.Tables(1, 8)
    nx, ny, nz = 48, 4, 4
    u, alpha, d = _drifting_pulse(O, T, nx, ny, nz, 1.5)
    ng = u.shape[0]
    zero = np.zeros(ng)
    x = -1.0 + (np.arange(nx) + 0.5) * d[0]
    m0 = T.M[0, 0]
    cJ = -1.0 / math.sqrt(3.0 / (4 * math.pi))

    def moments(uu):
        U = _interior(uu, nx, ny, nz)
        rho = U[..., 0].sum(axis=(1, 2)) / m0
        J = (cJ * U[..., 3]).sum()
        return rho, J

    rho0, J0 = moments(u)
    a = alpha.copy()
    uu = u.copy()
    dt = 0.3 * d[0]
    rc = _update(O, T, nx, ny, nz, uu, a, zero, zero, d, dt, solver=1, theta=2.0)
    assert rc == 0
    rho1, J1 = moments(uu)
    mass = rho0.sum()
    assert abs(rho1.sum() - mass) <= 1e-12 * mass  # sigma_a = 0: mass conserved
    dX = (x * rho1).sum() / mass - (x * rho0).sum() / mass
    want = dt * 0.5 * (J0 + J1) / mass
    assert J0 > 0 and abs(dX - want) <= 0.02 * abs(want), (dX, want)


def test_mn_step_is_the_reference_call_sequence(O):
    """orc_update (M_N, batched closure) equals update.cpp:185-275 composed from the pinned
    pieces: uOld <- u on the interior before anything else (206-212); per stage the ghost
    fill (218, comm.cpp:277-323), the closure over every cell including ghosts, which writes
    back both alpha and u (update.cpp:105-178), the kinetic flux of alpha (230), the Euler
    stage on u (eulerStep, 21-37); then u <- (u + uOld)/2 on the interior (268-274)."""
    T = O.Tables(1, 8)
    L = O.lib()
    nx, ny, nz = 6, 5, 4
    u, _, _ = O.init_grid(nx, ny, nz, T.nm, ic=1, sigma=0.3, floor=1e-2, noise=0.2, seed=5)
    rng = np.random.default_rng(5)
    u[:, 1:] = u[:, :1] * rng.uniform(-0.2, 0.2, (u.shape[0], T.nm - 1))
    alpha = O.init_alpha(u)
    ng = u.shape[0]
    sT = rng.uniform(0.5, 1.5, ng)
    sS = sT.copy()
    d, dt, theta = (2 / nx, 2 / ny, 2 / nz), 0.02, 1.0
    got_u, got_a = u.copy(), alpha.copy()
    assert _update(O, T, nx, ny, nz, got_u, got_a, sT, sS, d, dt, 1, theta=theta) == 0
    cu, ca = u.copy(), alpha.copy()
    uold = _interior(cu, nx, ny, nz).copy()
    for _ in range(2):
        L.orc_ghost_fill(nx, ny, nz, T.nm, cu)
        rc, _ = O.closure(T, cu, ca, O.opts())
        assert rc == 0
        F = np.zeros_like(cu)
        L.orc_flux_kinetic(nx, ny, nz, T.nm, T.nq, ca, F, O.C.byref(T.c), d[0], d[1], d[2],
                           theta, 0, np.zeros(ng * T.nq), 0)
        L.orc_euler(nx, ny, nz, T.nm, cu, F, sT, sS, dt)
    _interior(cu, nx, ny, nz)[...] = 0.5 * (_interior(cu, nx, ny, nz) + uold)
    assert np.array_equal(got_u, cu) and np.array_equal(got_a, ca)


def test_euler_stage_collision_closed_form(O):
    """eulerStep on zero flux is the collision term alone: u0 scaled by 1 + dt(sigma_s -
    sigma_t), higher moments by 1 - dt sigma_t; a unit flux enters with -dt (the divergence
    form u_t + div F = C u); ghosts untouched."""
    T = O.Tables(1, 4)
    nx, ny, nz = 3, 4, 5
    ng = (nx + 4) * (ny + 4) * (nz + 4)
    rng = np.random.default_rng(3)
    u = rng.uniform(1, 2, (ng, T.nm))
    sT = rng.uniform(0.5, 1.5, ng)
    sS = rng.uniform(0.0, 0.5, ng)
    dt = 0.1
    for fval in (0.0, 1.0):
        v = u.copy()
        O.lib().orc_euler(nx, ny, nz, T.nm, v, np.full_like(u, fval), sT, sS, dt)
        expect = u.copy()
        expect[:, 0] *= 1 + dt * (sS - sT)
        expect[:, 1:] *= (1 - dt * sT)[:, None]
        expect -= dt * fval
        vi, ei = _interior(v, nx, ny, nz), _interior(expect, nx, ny, nz)
        assert np.allclose(vi, ei, rtol=1e-15, atol=1e-15)
        ghost = np.ones((nx + 4, ny + 4, nz + 4), bool)
        ghost[2:-2, 2:-2, 2:-2] = False
        assert np.array_equal(v.reshape(nx + 4, ny + 4, nz + 4, -1)[ghost],
                              u.reshape(nx + 4, ny + 4, nz + 4, -1)[ghost])


# ------------------------------------------------------------ unbatched closure (opt.cpp)

@pytest.mark.parametrize("tol", [1e-8, 1e-3])
def test_unbatched_stops_on_the_relative_tolerance(O, tol):
    """optScaled's stop (opt.cpp:195, 298-300): ||g||_2 <= ||u_scaled||_2 tol and gamma <= 1.1,
    g recomputed here from the returned alpha and u by numpy.  At tol = 1e-3 some cells stop
    with ||g|| above tol itself, which only the relative rule allows."""
    T = O.Tables(2, 8)
    _, u, a0 = O.closure_batch_input(T, 400, 31, 0.5, 4.0)
    uu, aa = u.copy(), a0.copy()
    rc, st = O.closure(T, uu, aa, O.opts(tol=tol, unbatched=True, tolGamma=1.1))
    assert rc == 0
    e = np.exp(aa @ T.M.T) * T.w
    g = (e @ T.M - uu) / uu[:, :1]
    gn = np.linalg.norm(g, axis=1)
    rtol = tol * np.linalg.norm(uu / uu[:, :1], axis=1)
    assert (gn <= rtol * (1 + 1e-6) + 1e-14).all()
    assert np.allclose(st["err"], gn, rtol=1e-5, atol=1e-13)
    assert (st["gamma"] <= 1.1).all()
    if tol == 1e-3:
        assert (st["err"] > tol).sum() >= 5
    # the batched rule on the same cells is absolute: never above tol
    ub, ab = u.copy(), a0.copy()
    _, sb = O.closure(T, ub, ab, O.opts(tol=tol))
    assert (sb["err"] <= tol).all()


def test_unbatched_last_rung_gets_ten_times_max_iter(O):
    """opt.cpp:264-276: rungs 0-3 end after maxIter = 1 iteration each; entering the last
    rung multiplies maxIter by ten, so cells converge there instead of exhausting it."""
    T = O.Tables(2, 8)
    _, u, a0 = O.closure_batch_input(T, 200, 7, 2.0, 6.0)
    uu, aa = u.copy(), a0.copy()
    rc, st = O.closure(T, uu, aa, O.opts(maxIter=1, unbatched=True))
    assert rc == 0 and (st["flags"] & O_mod.FLAG_LADDER == 0).all()
    assert (st["regIndex"] == 4).all() and (st["r"] == 0.5).all()
    assert (st["totalIter"] >= 5).all() and (st["totalIter"] <= 14).all()
    assert (st["totalIter"] > 5).any()


@pytest.mark.skipif(not O_mod.Ref.available, reason="reference build absent")
@pytest.mark.parametrize("tol", [1e-8, 1e-4])
def test_unbatched_replay_with_reference_parts(O, tol):
    """optScaled (opt.cpp:139-324) re-driven from the reference's compiled fobjMn (with H)
    and linesearch, scipy's LAPACK dpotrf/dpocon/dpotrs and numpy's 2-norm, around opt()'s
    rescale (opt.cpp:330-355): the oracle's unbatched outcome tuples agree, and alpha to the
    factorisation's rounding."""
    T = O.Tables(2, 8)
    R = O.Ref(T)
    _, u, a0 = O.closure_batch_input(T, 60, 77, 0.5, 5.0)
    uo, ao = u.copy(), a0.copy()
    rc, so = O.closure(T, uo, ao, O.opts(tol=tol, unbatched=True))
    assert rc == 0
    regs = O_mod.REGULARIZATIONS
    p = int(math.floor(math.sqrt(T.nm))) + 1
    minf = math.sqrt((2 * p + 1) / (2 * math.pi))
    agree = 0
    for c in range(len(u)):
        u0 = u[c, 0]
        sc = math.log(u0) * 2.0 * math.sqrt(math.pi)
        a = a0[c].copy()
        a[0] -= sc
        uu = u[c] / u0
        alpha0, uorig = a.copy(), uu.copy()
        maxit = 100
        rtol = np.linalg.norm(uu) * tol
        it = tot = ig = reg = 0
        t = 1.0
        h = np.zeros((T.nm, T.nm))
        while True:
            it += 1
            tot += 1
            g = np.zeros(T.nm)
            f = R.L.ref_fobjMn(a, uu, g.ctypes.data, h.ctypes.data, 0, 0)
            ch, info = la.dpotrf(h, lower=1)
            cond = 1.0 / la.dpocon(ch, np.abs(h).sum(axis=1).max(), uplo="L")[0] if info == 0 else 0
            if tot == 1 and (info != 0 or cond > 1e10):
                alpha0 = np.zeros(T.nm)
                alpha0[0] = 2.0 * math.sqrt(math.pi) * math.log(0.5 / math.sqrt(math.pi))
                rtol = np.linalg.norm(uu) * tol
                a, it, ig, t = alpha0.copy(), 0, 0, 1.0
                continue
            if it >= maxit or info != 0 or t == 0 or cond > 1e10:
                assert reg < 4
                reg += 1
                uu = uorig * (1.0 - regs[reg])
                uu[0] = 1.0
                rtol = np.linalg.norm(uu) * tol
                a, ig, it, t = alpha0.copy(), 0, 0, 1.0
                if reg == 4:
                    maxit *= 10
                continue
            dd = -la.dpotrs(ch, g, lower=1)[0]
            err = np.linalg.norm(g)
            gam = math.exp(2.0 * np.abs(dd).sum() * minf)
            if err <= rtol and gam <= 1.1:
                break
            if err <= rtol:
                ig += 1
            t = R.L.ref_linesearch(a, np.ascontiguousarray(dd), f, g, uu)
        a[0] += sc
        same = (tot, ig, reg) == (so["totalIter"][c], so["iterGamma"][c], so["regIndex"][c])
        agree += same
        if same:
            assert np.abs(a - ao[c]).max() <= 1e-9 * np.abs(ao[c]).max()
    assert agree >= 57
