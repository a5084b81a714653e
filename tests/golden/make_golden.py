"""Writes tests/golden/golden.npz: reference values from numpy/scipy (present in this image)
that pin the oracle's table and LAPACK restatements.

  gl_Q{n}_x / _w     Gauss-Legendre nodes/weights, numpy.polynomial.legendre.leggauss
  plm                sqrt((2l+1)/4pi (l-m)!/(l+m)!) * scipy.special.lpmv(m, l, x) (lpmv
                     carries the Condon-Shortley phase, like GSL's sphPlm)
  pocon_H / _rcond   SPD matrices (row-major) and scipy.linalg.lapack.dpocon('L') rcond
                     with the max-row-sum norm, condition numbers 1e0 .. 1e12
  potrf_info         dpotrf('L') info for the same matrices plus indefinite ones

Run: python tests/golden/make_golden.py
"""
import math
import os

import numpy as np
import scipy.linalg.lapack as la
import scipy.special as sp

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = {}
    for Q in (1, 2, 3, 5, 8, 12, 24, 40):
        x, w = np.polynomial.legendre.leggauss(Q)
        out[f"gl_Q{Q}_x"] = x
        out[f"gl_Q{Q}_w"] = w
    xs = np.linspace(-0.999, 0.999, 37)
    rows = []
    for l in range(0, 9):
        for m in range(0, l + 1):
            norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) /
                             math.factorial(l + m))
            for x in xs:
                rows.append((l, m, x, norm * sp.lpmv(m, l, x)))
    out["plm"] = np.array(rows)
    rng = np.random.default_rng(20150101)
    Hs, rc, info = [], [], []
    for nm in (4, 9, 16, 25):
        for k in range(12):
            A = rng.normal(size=(nm, nm))
            Q_, _ = np.linalg.qr(A)
            s = np.logspace(0, -k, nm)
            H = (Q_ * s) @ Q_.T
            H = 0.5 * (H + H.T)
            Hp = np.zeros((25, 25))
            Hp[:nm, :nm] = H
            Hs.append(Hp)
            c, inf = la.dpotrf(H, lower=1)
            anorm = np.abs(H).sum(axis=1).max()
            r, inf2 = la.dpocon(c, anorm, uplo='L')
            rc.append((nm, r))
            info.append((nm, inf))
        H = np.eye(nm)
        H[nm // 2, nm // 2] = -1.0
        c, inf = la.dpotrf(H, lower=1)
        Hp = np.zeros((25, 25))
        Hp[:nm, :nm] = H
        Hs.append(Hp)
        rc.append((nm, np.nan))
        info.append((nm, inf))
    out["pocon_H"] = np.array(Hs)
    out["pocon_rcond"] = np.array(rc)
    out["potrf_info"] = np.array(info)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
