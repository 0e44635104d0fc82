"""Writes the golden vectors that pin the oracle's primitives.

Sources are textbook library routines (numpy.fft, numpy.linalg.eigh,
scipy.special.lpmv), not the reference: the reference has no tests or
fixtures of its own, and the vendor routines it calls (cuFFT, LAPACK dstev,
Matlab legendre) are not available here.

  fft_vectors.npz      seeded complex inputs and numpy.fft.fft outputs,
                       n in {24, 64, 96, 128, 192, 256}
  coriolis_eigh.npz    dense Coriolis matrices C_l (CoriolisMatrix.C:63-87
                       formula) and numpy eigh eigenvalues for a set of (J, l, n)
  legendre_scipy.npz   Matlab 'norm' associated Legendre values computed from
                       scipy.special.lpmv with the normalisation
                       sqrt((l+1/2)(l-m)!/(l+m)!) and the Condon-Shortley
                       phase removed

Run: python tests/golden/make_golden.py
"""
import math
import os

import numpy as np
from scipy.special import lpmv, gammaln

HERE = os.path.dirname(os.path.abspath(__file__))


def coriolis_dense(J, l, omega_min, n):
    c = np.zeros((n, n))
    for k in range(n):
        om = k + omega_min
        c[k, k] = J * (J + 1) - 2 * om * om
    lam = lambda a, om: math.sqrt(a * (a + 1.0) - om * (om + 1.0)) if a >= om else 0.0
    for k in range(n - 1):
        om = k + omega_min
        s = -lam(J, om) * lam(l, om) * math.sqrt(2.0 if om == 0 else 1.0)
        c[k, k + 1] = c[k + 1, k] = s
    return c


def main():
    rng = np.random.default_rng(20261203)
    fft = {}
    for n in (24, 64, 96, 128, 192, 256):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        fft[f"x{n}"] = x
        fft[f"f{n}"] = np.fft.fft(x)
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **fft)

    cases = [(1, 1, 0, 2), (4, 10, 1, 4), (20, 30, 0, 5), (20, 3, 0, 4), (8, 40, 0, 9), (40, 50, 1, 8), (3, 3, 1, 3)]
    mats, evals = [], []
    for J, l, om, n in cases:
        c = coriolis_dense(J, l, om, n)
        mats.append(c)
        evals.append(np.linalg.eigh(c)[0])
    np.savez_compressed(os.path.join(HERE, "coriolis_eigh.npz"), cases=np.array(cases),
                        **{f"c{i}": m for i, m in enumerate(mats)}, **{f"e{i}": e for i, e in enumerate(evals)})

    x = np.polynomial.legendre.leggauss(32)[0]
    vals = {}
    for m in (0, 1, 3, 8):
        for l in (m, m + 1, m + 5, 30):
            norm = math.exp(0.5 * (math.log(l + 0.5) + gammaln(l - m + 1) - gammaln(l + m + 1)))
            vals[f"m{m}_l{l}"] = (-1) ** m * norm * lpmv(m, l, x)   # lpmv includes (-1)^m
    np.savez_compressed(os.path.join(HERE, "legendre_scipy.npz"), x=x, **vals)


if __name__ == "__main__":
    main()
