"""Known-answer flux case: transmission through a symmetric Eckart barrier.

A J = 0 job whose potential separates, V(R, r, theta) = V_ho(R) + V_eck(r):
R is bound in the ground state of a harmonic oscillator, theta carries the
l = 0 state (the potential does not depend on theta, so no other l is ever
populated), and a Gaussian packet moves along r, the coordinate of the flux
surface, through the barrier
    V_eck(r) = V0 / cosh^2(alpha (r - r_b)).
The reaction probability at total energy E = E_R0 + E_trans is then the 1-D
transmission coefficient (Landau & Lifshitz, Quantum Mechanics, problem 4 to
section 25):
    T(E) = sinh^2(pi k / alpha) / (sinh^2(pi k / alpha) + cosh^2(pi sqrt(2 mu V0 - alpha^2/4) / alpha)),
    k = sqrt(2 mu E_trans),
provided eta_sq is |<phi_E|chi_0>|^2 of the r packet chi_0 at E_trans.
"""
import math

import numpy as np

from qd_openmp_cuda_amd import inputs

MU_R = 1000.0           # reduced mass along the flux coordinate (au)
MU_BIG_R = 20000.0      # mass along R
V0, ALPHA, R_B = 0.01, 2.0, 13.7
OMEGA_R = 0.001         # harmonic frequency along R (hartree)


def eckart_transmission(E_trans):
    E_trans = np.asarray(E_trans, dtype=np.float64)
    k = np.sqrt(2.0 * MU_R * E_trans)
    a = np.sinh(math.pi * k / ALPHA) ** 2
    b = np.cosh(math.pi * math.sqrt(2.0 * MU_R * V0 - ALPHA * ALPHA / 4.0) / ALPHA) ** 2
    return a / (a + b)


def eckart_job(n_energies=12, k_range=(3.6, 5.8), dt=5.0, total_steps=1600, n1=32, n2=512, ntheta=4,
               n_gradient_points=21):
    """Input dict (inputs.build_job layout) of the Eckart known-answer case."""
    R = inputs.radial_grid(1.9, 1.9 + 0.1 * (n1 - 1), n1, MU_BIG_R)
    r = inputs.radial_grid(0.2, 0.2 + 0.05 * (n2 - 1), n2, MU_R)
    x, w = inputs.gauss_legendre(ntheta)
    l_max = ntheta - 2
    omegas = np.array([0], dtype=np.int32)
    P = inputs.legendre_tables(omegas, l_max, x)
    Rc = 0.5 * (R["r"][0] + R["r"][-1])
    v_ho = 0.5 * MU_BIG_R * OMEGA_R ** 2 * (R["r"] - Rc) ** 2
    v_eck = V0 / np.cosh(ALPHA * (r["r"] - R_B)) ** 2
    V = np.asfortranarray(np.broadcast_to(v_ho[:, None, None] + v_eck[None, :, None], (n1, n2, ntheta)))
    # ground state of the oscillator along R, the incoming packet along r
    s = MU_BIG_R * OMEGA_R
    phi = (s / math.pi) ** 0.25 * np.exp(-0.5 * s * (R["r"] - Rc) ** 2)
    delta, r0, k0 = 1.0, 8.2, 4.7
    chi = (math.pi * delta * delta) ** -0.25 * np.exp(-(r["r"] - r0) ** 2 / (2 * delta * delta) + 1j * k0 * r["r"])
    Y = P[:, 0, 0]      # normalised P_0^0 at the nodes
    psi = np.zeros((n1, n2, ntheta, 1), dtype=np.complex128, order="F")
    psi[:, :, :, 0] = phi[:, None, None] * chi[None, :, None] * Y[None, None, :]
    # normalise on the grid (quadrature of the continuum-normalised factors)
    nrm = np.sum(np.abs(psi[:, :, :, 0]) ** 2 * w[None, None, :]) * R["dr"] * r["dr"]
    psi /= math.sqrt(nrm)
    d1 = np.ones(n1)
    # absorbers at both ends of the (periodic) r grid, each with a smooth inner
    # edge, so neither the transmitted nor the reflected part comes back
    d2 = inputs.cubic_absorber(r["r"], 8.0, 0.3) * inputs.cubic_absorber(-r["r"][::-1], 4.0, 0.2)[::-1]
    r_s = 15.7
    n_div = int(round((r_s - r["left"]) / r["dr"]))
    k = np.linspace(k_range[0], k_range[1], n_energies)
    e_trans = k * k / (2 * MU_R)
    e_r0 = 0.5 * OMEGA_R
    # |<phi_E|chi_0>|^2 for energy-normalised outgoing waves sqrt(mu/(2 pi k)) exp(i k r)
    eta = MU_R * delta / (k * math.sqrt(math.pi)) * np.exp(-delta * delta * (k - k0) ** 2)
    return {
        "r1": R, "r2": r,
        "theta": {"n": ntheta, "m": l_max, "x": x, "w": w},
        "pot": V,
        "time": {"total_steps": total_steps, "time_step": dt, "steps": 0},
        "dump1": d1, "dump2": d2,
        "omega_states": {"J": 0, "parity": 0, "l_max": l_max, "omegas": omegas,
                         "n_omegas_max": 1, "associated_legendres": P, "wave_packets": psi},
        "k_max": -1,
        "crp": {"energies": e_trans + e_r0, "eta_sq": eta, "n_dividing_surface": n_div,
                "n_gradient_points": n_gradient_points, "n_energies": n_energies, "calculate_CRP": 1},
        "meta": {"e_trans": e_trans, "k": k},
    }
