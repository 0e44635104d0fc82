"""Input construction for the propagator: grids, potential, Legendre tables,
initial wavepacket, absorbers and CRP normalisation.

This replaces the reference's Matlab driver (FH2main.m, InitWavePacket.m,
OmegaList.m, AssLegendreP.m), whose helper library (GaussLegendre2,
FH2PESJacobi, H2VibRotWaveFunction, EtaSq) is not part of the reference.
Every array is produced in the layout the reference's MEX entry reads
(column-major, first index fastest):

  pot                  (n1, n2, ntheta)            FH2main.m:147, cudaOpenMPEvolution.C:39-42
  associated_legendres (ntheta, nMax, n_omegas)    FH2main.m:41-50, matlabStructures.C:112-131
  wave_packets         (n1, n2, ntheta, n_omegas)  complex; FH2main.m:166-172

Conventions fixed here (the reference leaves them to un-vendored code):
  * Gauss-Legendre nodes in ascending order of x = cos(theta).
  * P_l^m is Matlab's legendre(l, x, 'norm'): sqrt((l+1/2)(l-m)!/(l+m)!) times
    the associated Legendre function WITHOUT the Condon-Shortley phase, so the
    leading coefficient is positive for every m; int_{-1}^{1} P^2 dx = 1.
  * Potential: LEPS surface with Sato parameter, capped at v_cap hartree.
  * Diatomic state: analytic Morse eigenfunction v of the H-H Morse curve.
  * Translational packet: g(R) = (pi d^2)^(-1/4) exp(-(R-R0)^2/(2 d^2) - i k0 R)
    (incoming, InitWavePacket.m:13-14), k0 = sqrt(2 mu_R E_coll).
  * Absorber: cubic, d(x) = exp(-strength ((x - x_a)/width)^3) for x > x_a.
"""
from __future__ import annotations

import dataclasses
import functools
import math
from typing import Optional, Sequence

import numpy as np

MASS_AU = 1822.888484929367          # amu -> electron masses (FH2main.m:64)
DEFAULT_MASSES = (19.00, 1.008, 1.008)


# ----------------------------------------------------------------------------
# grids and quadrature
# ----------------------------------------------------------------------------

def radial_grid(left: float, right: float, n: int, mass: float) -> dict:
    """linspace grid as the reference builds it (FH2main.m:82-84): dr = (right-left)/(n-1)."""
    r = np.linspace(left, right, n)
    return {"n": int(n), "left": float(r[0]), "dr": float(r[1] - r[0]), "mass": float(mass), "r": r}


def gauss_legendre(n: int):
    """Gauss-Legendre nodes (ascending) and weights on [-1, 1]."""
    x, w = np.polynomial.legendre.leggauss(n)
    return x.astype(np.float64), w.astype(np.float64)


def omega_list(J: int, p: int, l_max: int, k_max: Optional[int] = None) -> np.ndarray:
    """Helicity list (OmegaList.m:4-16), optionally truncated at k_max.

    Omega_min = 0 if (J + p) is even else 1; J = 0 forces p = 0 and Omega = {0}.
    """
    if J == 0:
        lo, hi = 0, 0
    else:
        lo = 0 if (J + p) % 2 == 0 else 1
        hi = min(J, l_max)
    if k_max is not None and k_max >= 0:
        hi = min(hi, k_max)
    if hi < lo:
        raise ValueError(f"empty helicity list for J={J} p={p} k_max={k_max}")
    return np.arange(lo, hi + 1, dtype=np.int32)


def assoc_legendre_norm(m: int, l_max: int, x: np.ndarray) -> np.ndarray:
    """Normalised associated Legendre functions P_l^m(x), l = m..l_max.

    Returns an array (len(x), l_max - m + 1) like AssLegendreP.m:9-12 (column
    l - m holds degree l).  Computed by the stable three-term recurrence on
    the normalised functions:
        P_m^m     = sqrt(1/2) prod_{i=1..m} sqrt((2i+1)/(2i)) (1-x^2)^{m/2}
        P_{m+1}^m = sqrt(2m+3) x P_m^m
        P_l^m     = a_l x P_{l-1}^m - b_l P_{l-2}^m
        a_l = sqrt((4l^2-1)/(l^2-m^2)),
        b_l = sqrt(((l-1)^2-m^2)(2l+1)/((2l-3)(l^2-m^2)))
    """
    x = np.asarray(x, dtype=np.float64)
    if l_max < m:
        raise ValueError("l_max < m")
    out = np.zeros((x.size, l_max - m + 1))
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    pmm = np.full_like(x, math.sqrt(0.5))
    for i in range(1, m + 1):
        pmm = pmm * math.sqrt((2 * i + 1) / (2 * i)) * s
    out[:, 0] = pmm
    if l_max == m:
        return out
    out[:, 1] = math.sqrt(2 * m + 3) * x * pmm
    for l in range(m + 2, l_max + 1):
        a = math.sqrt((4 * l * l - 1) / (l * l - m * m))
        b = math.sqrt(((l - 1) ** 2 - m * m) * (2 * l + 1) / ((2 * l - 3) * (l * l - m * m)))
        out[:, l - m] = a * x * out[:, l - m - 1] - b * out[:, l - m - 2]
    return out


def legendre_tables(omegas: np.ndarray, l_max: int, x: np.ndarray) -> np.ndarray:
    """associated_legendres (ntheta, nMax, n_omegas), FH2main.m:41-50."""
    n_max = l_max - int(omegas[0]) + 1
    P = np.zeros((x.size, n_max, len(omegas)), order="F")
    for b, om in enumerate(omegas):
        n = l_max - int(om) + 1
        P[:, :n, b] = assoc_legendre_norm(int(om), l_max, x)
    return P


# ----------------------------------------------------------------------------
# potential
# ----------------------------------------------------------------------------

@dataclasses.dataclass(frozen=True)
class LEPS:
    """London-Eyring-Polanyi-Sato surface for A + B2 (A = F, B = H).

    Pairs: (B-B) De, a, re and (A-B) De, a, re; Sato parameter k.
    Q and J integrals of the Sato-modified Morse / anti-Morse pair:
      Q = De/(4(1+k)) [(3+k) e^{-2a dx} - (2+6k) e^{-a dx}]
      J = De/(4(1+k)) [(1+3k) e^{-2a dx} - (6+2k) e^{-a dx}]
      V = Q1+Q2+Q3 - sqrt(((J1-J2)^2 + (J2-J3)^2 + (J3-J1)^2)/2)
    """
    de_bb: float = 0.174
    a_bb: float = 1.0
    re_bb: float = 1.40
    de_ab: float = 0.2
    a_ab: float = 1.1
    re_ab: float = 1.73
    sato: float = 0.1

    @staticmethod
    def _qj(r, de, a, re, k):
        e1 = np.exp(-a * (r - re))
        c = de / (4.0 * (1.0 + k))
        q = c * ((3.0 + k) * e1 * e1 - (2.0 + 6.0 * k) * e1)
        j = c * ((1.0 + 3.0 * k) * e1 * e1 - (6.0 + 2.0 * k) * e1)
        return q, j

    def energy(self, r_bb, r_ab1, r_ab2):
        q1, j1 = self._qj(r_bb, self.de_bb, self.a_bb, self.re_bb, self.sato)
        q2, j2 = self._qj(r_ab1, self.de_ab, self.a_ab, self.re_ab, self.sato)
        q3, j3 = self._qj(r_ab2, self.de_ab, self.a_ab, self.re_ab, self.sato)
        return q1 + q2 + q3 - np.sqrt(0.5 * ((j1 - j2) ** 2 + (j2 - j3) ** 2 + (j3 - j1) ** 2))

    def jacobi(self, R, r, x, v_cap: float = 0.5):
        """V(R_i, r_j, theta_k) on the Jacobi grid, (n1, n2, ntheta) Fortran order.

        r_AB = sqrt(R^2 + r^2/4 -/+ R r cos(theta)) for the two B atoms
        (equal masses).  Referenced to the B2 asymptote (-De_bb) and capped at
        v_cap hartree.
        """
        Rg = R[:, None, None]
        rg = r[None, :, None]
        xg = x[None, None, :]
        r1 = np.sqrt(np.maximum(Rg * Rg + 0.25 * rg * rg - Rg * rg * xg, 1e-12))
        r2 = np.sqrt(np.maximum(Rg * Rg + 0.25 * rg * rg + Rg * rg * xg, 1e-12))
        V = self.energy(np.broadcast_to(rg, r1.shape), r1, r2) + self.de_bb
        V = np.minimum(V, v_cap)
        return np.asfortranarray(V)


# ----------------------------------------------------------------------------
# initial state
# ----------------------------------------------------------------------------

def morse_state(r: np.ndarray, de: float, a: float, re: float, mu: float, v: int):
    """Analytic Morse eigenfunction v on the grid r, normalised sum(phi^2) dr = 1.

    lambda = sqrt(2 mu De)/a, z = 2 lambda exp(-a (r - re)),
    phi_v = z^{lambda - v - 1/2} exp(-z/2) L_v^{(2 lambda - 2 v - 1)}(z).
    Returns (phi, E_v) with E_v relative to the Morse minimum + De offset
    used by the LEPS reference (asymptote at 0): E_v = -De + w(v+1/2) - (w(v+1/2))^2/(4 De).
    """
    from scipy.special import eval_genlaguerre
    lam = math.sqrt(2.0 * mu * de) / a
    z = 2.0 * lam * np.exp(-a * (r - re))
    alpha = 2.0 * lam - 2.0 * v - 1.0
    logphi = (lam - v - 0.5) * np.log(z) - 0.5 * z
    phi = np.exp(logphi - logphi.max()) * eval_genlaguerre(v, alpha, z)
    dr = r[1] - r[0]
    phi = phi / math.sqrt(np.sum(phi * phi) * dr)
    w = a * math.sqrt(2.0 * de / mu)
    ev = w * (v + 0.5) - (w * (v + 0.5)) ** 2 / (4.0 * de)
    return phi, ev


def gaussian_packet(R: np.ndarray, R0: float, delta: float, k0: float) -> np.ndarray:
    """Incoming Gaussian (InitWavePacket.m:13-14)."""
    return (1.0 / (math.pi * delta * delta)) ** 0.25 * np.exp(
        -(R - R0) ** 2 / (2.0 * delta * delta) - 1j * k0 * R)


def eta_sq(E_coll: np.ndarray, mu: float, delta: float, k0: float) -> np.ndarray:
    """|<phi_E|g>|^2 for the Gaussian g and energy-normalised incoming waves
    phi_E = sqrt(mu/(2 pi k)) exp(-i k R):  mu d / (k sqrt(pi)) exp(-d^2 (k-k0)^2).
    Zero for E_coll <= 0."""
    E_coll = np.asarray(E_coll, dtype=np.float64)
    out = np.zeros_like(E_coll)
    pos = E_coll > 0
    k = np.sqrt(2.0 * mu * E_coll[pos])
    out[pos] = mu * delta / (k * math.sqrt(math.pi)) * np.exp(-delta * delta * (k - k0) ** 2)
    return out


def cubic_absorber(x: np.ndarray, width: float, strength: float) -> np.ndarray:
    """d(x) = exp(-strength ((x - x_a)/width)^3) on the last `width` bohr."""
    xa = x[-1] - width
    s = np.clip((x - xa) / width, 0.0, None)
    return np.exp(-strength * s ** 3)


def fornberg_first_derivative(n_points: int, h: float) -> np.ndarray:
    """Central finite-difference weights for d/dx on n_points (odd) equispaced
    points, exact for polynomials of degree n_points - 1:
    c_m = (-1)^(m+1) (k!)^2 / (m (k-m)! (k+m)!) / h, c_-m = -c_m, k = n_points // 2."""
    if n_points < 3 or n_points % 2 == 0:
        raise ValueError("n_points must be odd and >= 3")
    half = n_points // 2
    c = np.zeros(n_points)
    for m in range(1, half + 1):
        r = 1.0
        for q in range(m):
            r *= (half - q) / (half + q + 1)
        cm = (1.0 if m % 2 else -1.0) * r / m / h
        c[half + m] = cm
        c[half - m] = -cm
    return c


# ----------------------------------------------------------------------------
# whole job
# ----------------------------------------------------------------------------

@dataclasses.dataclass
class JobSpec:
    n1: int = 128
    n2: int = 64
    ntheta: int = 48
    R_range: tuple = (0.5, 14.0)
    r_range: tuple = (0.4, 8.0)
    l_max: Optional[int] = None        # default ntheta - 2 (largest the reference allows)
    J: int = 0
    parity: int = 0
    k_max: int = -1
    omega0: Optional[int] = None       # block holding the initial state (default omegas[0])
    v0: int = 0
    j0: int = 0
    R0: float = 9.0
    delta: float = 0.4
    E_coll: float = 0.05
    dt: float = 10.0
    total_steps: int = 2000
    masses: tuple = DEFAULT_MASSES
    leps: LEPS = dataclasses.field(default_factory=LEPS)
    v_cap: float = 0.5
    absorber_width: float = 2.0
    absorber_strength: float = 0.05
    absorber: bool = True
    flux_r: float = 4.5
    n_gradient_points: int = 21
    n_energies: int = 0
    energy_range: tuple = (0.01, 0.3)   # collision energies for CRP (hartree)
    seed: Optional[int] = None


@functools.lru_cache(maxsize=8)
def _potential_cached(leps, R_range, r_range, n1, n2, ntheta, v_cap):
    """The potential depends only on the surface and the grid: shared by every
    job of one shape (read-only)."""
    R = np.linspace(R_range[0], R_range[1], n1)
    r = np.linspace(r_range[0], r_range[1], n2)
    x, _ = gauss_legendre(ntheta)
    V = leps.jacobi(R, r, x, v_cap)
    V.flags.writeable = False
    return V


@functools.lru_cache(maxsize=64)
def _legendre_cached(omegas, l_max, ntheta):
    x, _ = gauss_legendre(ntheta)
    P = legendre_tables(np.array(omegas, dtype=np.int32), l_max, x)
    P.flags.writeable = False
    return P


def build_job(spec: JobSpec) -> dict:
    """Build every input array of one partial-wave job in the reference layouts."""
    mF, mH1, mH2 = (m * MASS_AU for m in spec.masses)
    mu_R = mF * (mH1 + mH2) / (mF + mH1 + mH2)           # FH2main.m:85-86
    mu_r = mH1 * mH2 / (mH1 + mH2)                       # FH2main.m:108
    r1 = radial_grid(spec.R_range[0], spec.R_range[1], spec.n1, mu_R)
    r2 = radial_grid(spec.r_range[0], spec.r_range[1], spec.n2, mu_r)
    x, w = gauss_legendre(spec.ntheta)
    l_max = spec.ntheta - 2 if spec.l_max is None else spec.l_max
    J, p = spec.J, (0 if spec.J == 0 else spec.parity)
    omegas = omega_list(J, p, l_max, spec.k_max if spec.k_max >= 0 else None)
    P = _legendre_cached(tuple(int(o) for o in omegas), l_max, spec.ntheta)
    V = _potential_cached(spec.leps, spec.R_range, spec.r_range, spec.n1, spec.n2, spec.ntheta, spec.v_cap)

    R0 = spec.R0
    if spec.seed is not None:
        rng = np.random.default_rng(spec.seed)
        R0 = float(rng.uniform(8.0, 10.0))
    k0 = math.sqrt(2.0 * mu_R * spec.E_coll)
    g = gaussian_packet(r1["r"], R0, spec.delta, k0)
    phi, e_vib = morse_state(r2["r"], spec.leps.de_bb, spec.leps.a_bb, spec.leps.re_bb, mu_r, spec.v0)
    omega0 = int(omegas[0]) if spec.omega0 is None else int(spec.omega0)
    if omega0 not in omegas or spec.j0 < omega0:
        raise ValueError("initial omega must be in the helicity list and <= j0")
    Pj = assoc_legendre_norm(omega0, spec.j0, x)[:, spec.j0 - omega0]
    psi = np.zeros((spec.n1, spec.n2, spec.ntheta, len(omegas)), dtype=np.complex128, order="F")
    psi[:, :, :, int(np.where(omegas == omega0)[0][0])] = (
        g[:, None, None] * phi[None, :, None] * Pj[None, None, :])
    # rotational energy of the diatomic (expectation of j(j+1)/(2 mu r^2) in phi)
    e_rot = spec.j0 * (spec.j0 + 1) * float(np.sum(phi * phi / (2 * mu_r * r2["r"] ** 2)) * r2["dr"])
    e_int = -spec.leps.de_bb + e_vib + e_rot + spec.leps.de_bb

    if spec.absorber:
        d1 = cubic_absorber(r1["r"], spec.absorber_width, spec.absorber_strength)
        d2 = cubic_absorber(r2["r"], spec.absorber_width, spec.absorber_strength)
    else:
        d1 = d2 = None

    n_div = int(round((spec.flux_r - r2["left"]) / r2["dr"]))
    if spec.n_energies > 0:
        e_coll = np.linspace(spec.energy_range[0], spec.energy_range[1], spec.n_energies)
        energies = e_coll + e_int
        esq = eta_sq(e_coll, mu_R, spec.delta, k0)
    else:
        energies = np.zeros(0)
        esq = np.zeros(0)

    return {
        "r1": r1, "r2": r2,
        "theta": {"n": spec.ntheta, "m": l_max, "x": x, "w": w},
        "pot": V,
        "time": {"total_steps": spec.total_steps, "time_step": spec.dt, "steps": 0},
        "dump1": d1, "dump2": d2,
        "omega_states": {"J": J, "parity": p, "l_max": l_max, "omegas": omegas,
                         "n_omegas_max": len(omegas), "associated_legendres": P,
                         "wave_packets": psi},
        "k_max": spec.k_max,
        "crp": {"energies": energies, "eta_sq": esq, "n_dividing_surface": n_div,
                "n_gradient_points": spec.n_gradient_points, "n_energies": len(energies),
                "calculate_CRP": 1 if len(energies) else 0},
        "meta": {"mu_R": mu_R, "mu_r": mu_r, "k0": k0, "R0": R0, "e_int": e_int,
                 "omega0": omega0, "E_coll": spec.E_coll, "delta": spec.delta},
    }


# ----------------------------------------------------------------------------
# the workloads
# ----------------------------------------------------------------------------

def config_spec(name: str, **over) -> JobSpec:
    """Job specs of the five workloads.

    c1: 128x64x48, J=0, dt=10, 2000 steps (the CPU-runnable case)
    c2: 256x192x96, J=0, R0 seeded in [8, 10], 20000 steps
    c3: 192x128x64, J=20, K_max=4 (5 omega blocks), 10000 steps
    c4: c2 in mixed precision with P(E) at 200 energies in [0.01, 0.3]
    """
    if name == "c1":
        s = JobSpec()
    elif name == "c2":
        s = JobSpec(n1=256, n2=192, ntheta=96, total_steps=20000, seed=2)
    elif name == "c3":
        s = JobSpec(n1=192, n2=128, ntheta=64, J=20, parity=0, k_max=4, total_steps=10000)
    elif name == "c4":
        s = JobSpec(n1=256, n2=192, ntheta=96, total_steps=20000, seed=2, n_energies=200)
    else:
        raise ValueError(name)
    return dataclasses.replace(s, **over)


def serve_jobs(n_jobs: int = 64, seed: int = 5, rate_per_s: float = 4.0,
               shapes: Sequence[tuple] = ((128, 64, 48), (192, 128, 64), (256, 192, 96)),
               j_max: int = 40, k_cap: int = 8, total_steps: int = 200, n_energies: int = 0):
    """Seeded Poisson stream of partial-wave jobs (J uniform in 0..j_max,
    K_max = min(J, k_cap), v0 in {0,1}, j0 in {0,1,2}, shape from `shapes`;
    reaction probabilities at n_energies collision energies in [0.01, 0.3]
    hartree when n_energies > 0).  Returns a list of (arrival_seconds, JobSpec)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for _ in range(n_jobs):
        t += float(rng.exponential(1.0 / rate_per_s))
        J = int(rng.integers(0, j_max + 1))
        n1, n2, nt = shapes[int(rng.integers(0, len(shapes)))]
        v0 = int(rng.integers(0, 2))
        j0 = int(rng.integers(0, 3))
        ec = float(rng.uniform(0.02, 0.2))
        p = 0 if J == 0 else int(rng.integers(0, 2))
        lo = 0 if (J == 0 or (J + p) % 2 == 0) else 1
        j0 = max(j0, lo)
        spec = JobSpec(n1=n1, n2=n2, ntheta=nt, J=J, parity=p, k_max=min(J, k_cap),
                       v0=v0, j0=j0, E_coll=ec, total_steps=total_steps, n_energies=n_energies)
        jobs.append((t, spec))
    return jobs


# ----------------------------------------------------------------------------
# fixture files for the command-line driver (driver/qd_fixture_run.cpp)
# ----------------------------------------------------------------------------

def write_fixture(job: dict, path: str, total_steps: Optional[int] = None) -> None:
    """Write one job in the binary fixture layout read by qd_fixture_run."""
    r1, r2, th, os_, crp = job["r1"], job["r2"], job["theta"], job["omega_states"], job["crp"]
    P = np.asfortranarray(os_["associated_legendres"], dtype=np.float64)
    omegas = np.asarray(os_["omegas"], dtype=np.int32)
    psi = np.asfortranarray(os_["wave_packets"], dtype=np.complex128)
    has_dump = job["dump1"] is not None and job["dump2"] is not None
    ts = job["time"]["total_steps"] if total_steps is None else total_steps
    header = np.array([r1["n"], r2["n"], th["n"], th["m"], len(omegas), os_["J"], os_["parity"], os_["l_max"],
                       P.shape[1], ts, job["time"]["steps"], job.get("k_max", -1), len(crp["energies"]),
                       crp["n_dividing_surface"], crp["n_gradient_points"], crp["calculate_CRP"],
                       int(has_dump)], dtype=np.int32)
    geom = np.array([r1["left"], r1["dr"], r1["mass"], r2["left"], r2["dr"], r2["mass"],
                     job["time"]["time_step"]], dtype=np.float64)
    with open(path, "wb") as f:
        f.write(b"QDFX0001")
        f.write(header.tobytes())
        f.write(geom.tobytes())
        f.write(np.ascontiguousarray(th["x"], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(th["w"], dtype=np.float64).tobytes())
        f.write(np.asfortranarray(job["pot"], dtype=np.float64).tobytes(order="F"))
        f.write(omegas.tobytes())
        f.write(P.tobytes(order="F"))
        f.write(psi.tobytes(order="F"))
        if has_dump:
            f.write(np.ascontiguousarray(job["dump1"], dtype=np.float64).tobytes())
            f.write(np.ascontiguousarray(job["dump2"], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(crp["energies"], dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(crp["eta_sq"], dtype=np.float64).tobytes())


def read_fixture_output(path: str, job: dict, total_steps: int) -> dict:
    """Read the outputs qd_fixture_run wrote for `job`."""
    r1, r2, th, os_ = job["r1"], job["r2"], job["theta"], job["omega_states"]
    nO = len(os_["omegas"])
    n_e = len(job["crp"]["energies"])
    shape = (r1["n"], r2["n"], th["n"], nO)
    with open(path, "rb") as f:
        if f.read(8) != b"QDOUT001":
            raise ValueError(f"{path} is not a driver output file")
        steps = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
        obs = np.frombuffer(f.read(8 * total_steps * nO * 5), dtype=np.float64).reshape(total_steps, nO, 5)
        n_psi = int(np.prod(shape))
        psi = np.frombuffer(f.read(16 * n_psi), dtype=np.complex128).reshape(shape, order="F")
        crp = np.frombuffer(f.read(8 * n_e), dtype=np.float64)
    return {"steps": steps, "obs": obs, "psi": psi, "crp": crp}
