"""Independent numpy/scipy statement of the split-operator step, for pinning
the oracle on small grids.

It shares no code with the oracle: the 2-D DFT is numpy.fft, the Coriolis
exponential is scipy.linalg.expm of the tridiagonal matrix written from
CoriolisMatrix.C:63-87, the Legendre transforms are numpy tensordots, and the
rotational phase is applied as a quarter-step factor on both sides of the
Coriolis exponential exactly as cudaOpenMP2.cu:350-354 / :364-368 order them.
"""
import math

import numpy as np
from scipy.linalg import expm


def coriolis_matrix(J, l, omega_min, n):
    c = np.zeros((n, n))
    lam = lambda a, om: math.sqrt(a * (a + 1.0) - om * (om + 1.0)) if a >= om else 0.0
    for k in range(n):
        om = k + omega_min
        c[k, k] = J * (J + 1) - 2 * om * om
    for k in range(n - 1):
        om = k + omega_min
        c[k, k + 1] = c[k + 1, k] = -lam(J, om) * lam(l, om) * math.sqrt(2.0 if om == 0 else 1.0)
    return c


class NumpyPropagator:
    def __init__(self, job, k_max=None):
        self.job = job
        r1, r2 = job["r1"], job["r2"]
        os_ = job["omega_states"]
        self.n1, self.n2, self.nt = r1["n"], r2["n"], job["theta"]["n"]
        self.R = r1["left"] + r1["dr"] * np.arange(self.n1)
        self.r = r2["left"] + r2["dr"] * np.arange(self.n2)
        self.m1, self.m2 = r1["mass"], r2["mass"]
        self.dr1, self.dr2 = r1["dr"], r2["dr"]
        self.w = job["theta"]["w"]
        self.V = np.asarray(job["pot"])
        self.omegas = [int(o) for o in os_["omegas"]]
        self.lmax = os_["l_max"]
        self.J = os_["J"]
        self.kmax = self.omegas[-1]
        self.P = [os_["associated_legendres"][:, : self.lmax - om + 1, b] for b, om in enumerate(self.omegas)]
        self.psi = [np.array(os_["wave_packets"][:, :, :, b]) for b in range(len(self.omegas))]
        self.dt = job["time"]["time_step"]
        self.steps = job["time"]["steps"]
        k1 = 2 * np.pi * np.fft.fftfreq(self.n1, d=self.dr1)
        k2 = 2 * np.pi * np.fft.fftfreq(self.n2, d=self.dr2)
        # fftfreq puts n/2 at -n/2; the reference's table uses +n/2 (same square)
        self.T = k1[:, None] ** 2 / (2 * self.m1) + k2[None, :] ** 2 / (2 * self.m2)
        self.Irot = 1 / (2 * self.m1 * self.R[:, None] ** 2) + 1 / (2 * self.m2 * self.r[None, :] ** 2)
        self.d = None
        if job["dump1"] is not None and job["dump2"] is not None:
            self.d = np.asarray(job["dump1"])[:, None] * np.asarray(job["dump2"])[None, :]

    def _rot(self, F, tau):
        for b, om in enumerate(self.omegas):
            for l in range(om, self.lmax + 1):
                F[b][:, :, l] *= np.exp(-1j * tau * l * (l + 1) * self.Irot)

    def _cor(self, F, tau, energy):
        nO = len(self.omegas)
        om0 = self.omegas[0]
        ce = np.zeros(nO)
        if energy:
            for l in range(om0, self.lmax + 1):
                n = min(self.kmax, l) - om0 + 1
                C = coriolis_matrix(self.J, l, om0, n)
                for b in range(n):
                    for b1 in (b, b + 1):
                        if b1 >= n:
                            continue
                        c = C[b, b1] / (2 * self.m1 * self.R ** 2)
                        dot = np.sum(np.conj(F[b][:, :, l]) * c[:, None] * F[b1][:, :, l])
                        ce[b] += (2 if b1 != b else 1) * dot.real * self.dr1 * self.dr2
        for l in range(om0, self.lmax + 1):
            n = min(self.kmax, l) - om0 + 1
            C = coriolis_matrix(self.J, l, om0, n)
            X = np.stack([F[b][:, :, l] for b in range(n)])      # (n, n1, n2)
            Y = np.zeros_like(X)
            for i in range(self.n1):
                E = expm(-1j * tau * C / (2 * self.m1 * self.R[i] ** 2))
                Y[:, i, :] = E @ X[:, i, :]
            for b in range(n):
                F[b][:, :, l] = Y[b]
        return ce

    def step(self):
        dt = self.dt
        nO = len(self.omegas)
        L = self.lmax + 1
        obs = np.zeros((nO, 5))
        for b in range(nO):
            if self.d is not None:
                self.psi[b] = self.psi[b] * self.d[:, :, None]
            if self.steps == 0:
                self.psi[b] = self.psi[b] * np.exp(1j * dt / 2 * self.V)
            self.psi[b] = self.psi[b] * np.exp(-1j * dt * self.V)
        F = []
        for b, om in enumerate(self.omegas):
            f = np.zeros((self.n1, self.n2, L), dtype=complex)
            f[:, :, om:] = np.tensordot(self.psi[b], self.w[:, None] * self.P[b], axes=([2], [0]))
            F.append(f)
        self._rot(F, dt / 4)
        self._cor(F, dt / 2, False)
        self._rot(F, dt / 4)
        for b, om in enumerate(self.omegas):
            g = np.fft.fft2(F[b][:, :, om:], axes=(0, 1))
            g *= np.exp(-1j * dt * self.T)[:, :, None]
            obs[b, 2] = np.sum(np.abs(g) ** 2 * self.T[:, :, None]) * self.dr1 * self.dr2 / (self.n1 * self.n2)
            F[b][:, :, om:] = np.fft.ifft2(g, axes=(0, 1))
        self._rot(F, dt / 4)
        obs[:, 4] = self._cor(F, dt / 2, True)
        self._rot(F, dt / 4)
        for b, om in enumerate(self.omegas):
            ls = np.arange(om, L)
            obs[b, 3] = np.sum(np.abs(F[b][:, :, om:]) ** 2 * self.Irot[:, :, None] * (ls * (ls + 1))[None, None, :]) \
                * self.dr1 * self.dr2
            self.psi[b] = np.tensordot(F[b][:, :, om:], self.P[b].T, axes=([2], [0]))
            q = np.abs(self.psi[b]) ** 2
            obs[b, 0] = np.sum(q * self.w[None, None, :]) * self.dr1 * self.dr2
            obs[b, 1] = np.sum(q * self.V * self.w[None, None, :]) * self.dr1 * self.dr2
        self.steps += 1
        return obs

    def state(self):
        return np.stack(self.psi, axis=3)
