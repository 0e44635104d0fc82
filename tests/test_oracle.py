"""Pins of the FP64 CPU oracle (oracle/oracle.cpp) — CPU only.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails: golden vectors from
textbook library routines, closed forms, invariants and an independent
numpy/scipy restatement of the step.
"""
import dataclasses
import math
import os

import numpy as np
import pytest

from qd_openmp_cuda_amd import abi, inputs
from oracle import oracle
from numpy_step import NumpyPropagator, coriolis_matrix

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 8, 12, 15, 16, 30, 32])
def test_fft_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    assert np.allclose(oracle.fft1d(x, -1), W @ x, atol=1e-12, rtol=0)
    assert np.allclose(oracle.fft1d(x, +1), np.conj(W) @ x, atol=1e-12, rtol=0)


def test_fft_golden_vectors():
    g = np.load(os.path.join(GOLD, "fft_vectors.npz"))
    for n in (24, 64, 96, 128, 192, 256):
        x, f = g[f"x{n}"], g[f"f{n}"]
        y = oracle.fft1d(x, -1)
        assert np.abs(y - f).max() < 1e-12 * np.abs(f).max()
        # Parseval and inverse
        assert abs(np.vdot(y, y).real - n * np.vdot(x, x).real) < 1e-10 * n * np.vdot(x, x).real
        assert np.abs(oracle.fft1d(y, +1) / n - x).max() < 1e-13


# ---------------------------------------------------------------- Coriolis
def test_coriolis_matrix_closed_form():
    # J = l = 1, omega_min = 0: diag (2, 0), off = -sqrt2*sqrt2*sqrt2 = -2 sqrt2
    c = oracle.coriolis_dense(1, 1, 0, 2)
    assert np.allclose(c, [[2.0, -2 * math.sqrt(2)], [-2 * math.sqrt(2), 0.0]], atol=1e-14)
    # omega >= 1: no sqrt2 factor; J=3, l=2, omega 1..2: off = -sqrt(12-2)*sqrt(6-2)
    c = oracle.coriolis_dense(3, 2, 1, 2)
    assert np.allclose(c, [[12 - 2, -math.sqrt(10) * 2], [-math.sqrt(10) * 2, 12 - 8]], atol=1e-13)


def test_coriolis_eigen_golden():
    g = np.load(os.path.join(GOLD, "coriolis_eigh.npz"))
    for i, (J, l, om, n) in enumerate(g["cases"]):
        c = oracle.coriolis_dense(int(J), int(l), int(om), int(n))
        assert np.allclose(c, g[f"c{i}"], atol=1e-12)
        e, v = oracle.sym_eig(c)
        assert np.allclose(np.sort(e), g[f"e{i}"], atol=1e-10 * max(1, np.abs(e).max()))
        assert np.allclose(v @ np.diag(e) @ v.T, c, atol=1e-10 * max(1, np.abs(c).max()))
        assert np.allclose(v.T @ v, np.eye(int(n)), atol=1e-13)


def test_coriolis_exponential_closed_form_and_expm():
    from scipy.linalg import expm
    # n = 2: exp(-i s C) for symmetric C = a I + (d sigma_z + o sigma_x):
    # exp(-i s a) [cos(s w) I - i sin(s w) (d sigma_z + o sigma_x)/w], w = sqrt(d^2 + o^2)
    J, l, tau, I = 1, 3, 5.0, 37.0
    C = oracle.coriolis_dense(J, l, 0, 2)
    a, d, o = (C[0, 0] + C[1, 1]) / 2, (C[0, 0] - C[1, 1]) / 2, C[0, 1]
    w = math.hypot(d, o)
    s = tau / I
    closed = np.exp(-1j * s * a) * (math.cos(s * w) * np.eye(2)
                                    - 1j * math.sin(s * w) * np.array([[d, o], [o, -d]]) / w)
    E = oracle.coriolis_exp(J, l, 0, 2, tau, I)
    assert np.abs(E - closed).max() < 1e-13
    for (J, l, om, n) in [(20, 30, 0, 5), (8, 40, 1, 8), (4, 2, 0, 3)]:
        E = oracle.coriolis_exp(J, l, om, n, 5.0, 123.0)
        ref = expm(-1j * 5.0 / 123.0 * coriolis_matrix(J, l, om, n))
        assert np.abs(E - ref).max() < 1e-12
        assert np.abs(E.conj().T @ E - np.eye(n)).max() < 1e-13     # unitary


def test_coriolis_j0_identity():
    E = oracle.coriolis_exp(0, 5, 0, 1, 10.0, 50.0)
    assert abs(E[0, 0] - 1.0) < 1e-15


# ---------------------------------------------------------------- Legendre tables (input side)
def test_legendre_tables_match_scipy_and_are_orthonormal():
    g = np.load(os.path.join(GOLD, "legendre_scipy.npz"))
    x = g["x"]
    for m in (0, 1, 3, 8):
        P = inputs.assoc_legendre_norm(m, 30, x)
        for l in (m, m + 1, m + 5, 30):
            assert np.abs(P[:, l - m] - g[f"m{m}_l{l}"]).max() < 1e-12
    xg, wg = inputs.gauss_legendre(48)
    for m in (0, 2, 7):
        P = inputs.assoc_legendre_norm(m, 46, xg)
        assert np.abs(P.T @ (wg[:, None] * P) - np.eye(P.shape[1])).max() < 1e-13


# ---------------------------------------------------------------- whole step
def small_spec(**kw):
    base = dict(n1=32, n2=24, ntheta=12, R_range=(2.0, 10.0), r_range=(0.6, 4.0), E_coll=0.01,
                absorber=False, dt=10.0)
    base.update(kw)
    return inputs.JobSpec(**base)


def run_oracle(job, steps, start=0, threads=4):
    inp = abi.Inputs(job, total_steps=steps, steps=start)
    oracle.evolve(inp, threads)
    return inp


@pytest.mark.parametrize("kw", [
    dict(),
    dict(absorber=True, absorber_width=1.5, absorber_strength=0.1),
    dict(J=3, parity=1, k_max=-1, omega0=1, j0=1),
    dict(J=6, parity=0, k_max=2, omega0=0, j0=2),
])
def test_oracle_matches_numpy_restatement(kw):
    job = inputs.build_job(small_spec(**kw))
    ref = NumpyPropagator(job)
    obs = [ref.step() for _ in range(3)]
    o = run_oracle(job, 3)
    e = np.linalg.norm(o.psi - ref.state()) / np.linalg.norm(ref.state())
    assert e < 1e-12
    assert np.abs(o.obs - np.array(obs)).max() < 1e-12
    assert o.steps[0] == 3


def test_free_particle_exact_propagator():
    """V = 0, J = 0, l = 0 state: the step reduces to the exact FFT propagator
    psi(t) = IFFT2(exp(-i t T) FFT2(psi0)) on every theta slice."""
    job = inputs.build_job(small_spec())
    job["pot"] = np.zeros_like(job["pot"])
    n = 5
    o = run_oracle(job, n)
    p0 = job["omega_states"]["wave_packets"][:, :, :, 0]
    r1, r2 = job["r1"], job["r2"]
    k1 = 2 * np.pi * np.fft.fftfreq(r1["n"], d=r1["dr"])
    k2 = 2 * np.pi * np.fft.fftfreq(r2["n"], d=r2["dr"])
    T = k1[:, None] ** 2 / (2 * r1["mass"]) + k2[None, :] ** 2 / (2 * r2["mass"])
    t = n * job["time"]["time_step"]
    exact = np.fft.ifft2(np.fft.fft2(p0, axes=(0, 1)) * np.exp(-1j * t * T)[:, :, None], axes=(0, 1))
    assert np.abs(o.psi[:, :, :, 0] - exact).max() < 1e-12 * np.abs(exact).max()
    # kinetic energy equals <T> of the initial state exactly (free motion)
    ke0 = np.sum(np.abs(np.fft.fft2(p0, axes=(0, 1))) ** 2 * T[:, :, None] * job["theta"]["w"][None, None, :]) \
        * r1["dr"] * r2["dr"] / (r1["n"] * r2["n"])
    assert np.abs(o.obs[:, 0, 2] - ke0).max() < 1e-12 * ke0


def _free_acf_closed_form(job, times):
    """<psi_0|psi(t)> for V = 0 and an l = 0 state: psi(t) = IFFT2(exp(-i t T) FFT2 psi_0)
    on every theta slice, summed with the quadrature weights and dr1 dr2."""
    p0 = job["omega_states"]["wave_packets"][:, :, :, 0]
    r1, r2 = job["r1"], job["r2"]
    k1 = 2 * np.pi * np.fft.fftfreq(r1["n"], d=r1["dr"])
    k2 = 2 * np.pi * np.fft.fftfreq(r2["n"], d=r2["dr"])
    T = k1[:, None] ** 2 / (2 * r1["mass"]) + k2[None, :] ** 2 / (2 * r2["mass"])
    f0 = np.fft.fft2(p0, axes=(0, 1))
    w = job["theta"]["w"]
    out = []
    for t in times:
        pt = np.fft.ifft2(f0 * np.exp(-1j * t * T)[:, :, None], axes=(0, 1))
        out.append(np.sum(np.conj(p0) * pt * w[None, None, :]) * r1["dr"] * r2["dr"])
    return np.array(out)


def test_autocorrelation_free_particle_closed_form():
    """The oracle's autocorrelation, record n at t_n = (n + 1) dt, against the
    closed form <psi_0|IFFT2(exp(-i t T) FFT2 psi_0)>.  Its imaginary part is
    as large as its real part here, so a conjugate on the wrong side fails, and
    dropping the weights w_k (sum 2) or dr1 dr2 changes the modulus."""
    job = inputs.build_job(small_spec(E_coll=0.02))
    job["pot"] = np.zeros_like(job["pot"])
    n, dt = 8, job["time"]["time_step"]
    o = run_oracle(job, n)
    exact = _free_acf_closed_form(job, dt * np.arange(1, n + 1))
    assert np.abs(exact.imag).max() > 0.3 * np.abs(exact).max()
    assert np.abs(exact - 1).max() > 1e-2          # the packet has moved
    assert np.abs(o.acf - exact).max() < 1e-12


def test_autocorrelation_constant_potential_phase():
    """A constant potential V0 commutes with every factor of the step, so the
    autocorrelation is exp(-i V0 t_n) times the V = 0 one.  The stored state
    carries exp(-i V0 (t_n - dt/2)) (the leapfrogged half step): an
    autocorrelation taken on it without restoring V(dt/2) is off by the phase
    V0 dt / 2 = 0.05 here."""
    job = inputs.build_job(small_spec(E_coll=0.02))
    V0, n, dt = 0.01, 8, job["time"]["time_step"]
    job["pot"] = np.zeros_like(job["pot"])
    free = run_oracle(job, n).acf
    job["pot"] = np.full_like(job["pot"], V0)
    shifted = run_oracle(job, n).acf
    t = dt * np.arange(1, n + 1)
    assert np.abs(shifted - np.exp(-1j * V0 * t) * free).max() < 1e-12
    assert np.abs(shifted - np.exp(-1j * V0 * (t - dt / 2)) * free).max() > 1e-2


def test_autocorrelation_across_resume():
    """Resumed at steps = N1 from the stored state, the autocorrelation refers
    to the state at t = N1 dt; with V = V0 constant and no absorber the
    evolution is a group, so record m equals record m of a fresh run."""
    job = inputs.build_job(small_spec(E_coll=0.02))
    job["pot"] = np.full_like(job["pot"], 0.01)
    n1, n2 = 3, 6
    fresh = run_oracle(job, n2)
    first = run_oracle(job, n1)
    res = dict(job)
    res["omega_states"] = dict(job["omega_states"], wave_packets=first.psi)
    second = run_oracle(res, n2, start=n1)
    assert second.steps[0] == n1 + n2
    assert np.abs(second.acf - fresh.acf).max() < 1e-12


def test_rigid_rotor_single_l():
    """V = 0 and a single-l angular state: the rotational energy is l(l+1)<I1+I2>
    of the radial density at every step and the norm is conserved."""
    spec = small_spec(j0=3)
    job = inputs.build_job(spec)
    job["pot"] = np.zeros_like(job["pot"])
    o = run_oracle(job, 4)
    # closed form on the stored state after each step: recompute from o.psi at the end
    r1, r2 = job["r1"], job["r2"]
    R = r1["left"] + r1["dr"] * np.arange(r1["n"])
    r = r2["left"] + r2["dr"] * np.arange(r2["n"])
    Irot = 1 / (2 * r1["mass"] * R[:, None] ** 2) + 1 / (2 * r2["mass"] * r[None, :] ** 2)
    dens = np.sum(np.abs(o.psi[:, :, :, 0]) ** 2 * job["theta"]["w"][None, None, :], axis=2)
    assert abs(o.obs[-1, 0, 3] - 12 * np.sum(dens * Irot) * r1["dr"] * r2["dr"]) < 1e-12
    assert np.abs(o.obs[:, 0, 0] - o.obs[0, 0, 0]).max() < 1e-13


def test_time_reversal_and_norm():
    """theta-independent V, no absorber: the Legendre projection commutes with V,
    every operator is unitary, and running -dt after +dt returns psi0."""
    spec = small_spec(J=4, parity=0, k_max=-1, omega0=0, j0=2)
    job = inputs.build_job(spec)
    V = job["pot"]
    job["pot"] = np.asfortranarray(np.repeat(V[:, :, :1], V.shape[2], axis=2))
    n = 6
    fwd = run_oracle(job, n)
    norm = fwd.obs[:, :, 0].sum(1)
    assert np.abs(norm - norm[0]).max() < 1e-13
    # undo: the stored state lacks V(dt/2); reverse run starts at steps=0 with -dt
    psi_half = fwd.psi * np.exp(-1j * 5.0 * job["pot"])[:, :, :, None]      # apply V(dt/2)
    back = dict(job)
    back["time"] = dict(job["time"], time_step=-10.0)
    back["omega_states"] = dict(job["omega_states"], wave_packets=psi_half)
    rev = run_oracle(back, n)
    psi_rev = rev.psi * np.exp(1j * 5.0 * job["pot"])[:, :, :, None]       # V(-dt/2)
    p0 = job["omega_states"]["wave_packets"]
    assert np.linalg.norm(psi_rev - p0) / np.linalg.norm(p0) < 1e-12


def test_energy_conservation_small_dt():
    spec = small_spec(J=4, parity=0, k_max=2, omega0=0, j0=1, dt=1.0)
    job = inputs.build_job(spec)
    o = run_oracle(job, 40)
    E = o.obs[:, :, 1:].sum(axis=(1, 2))
    assert np.abs(E - E[0]).max() < 2e-4 * abs(E[0])


def test_norm_check_and_copy_back():
    job = inputs.build_job(small_spec())
    calls = []
    inp = abi.Inputs(job, total_steps=5, copy_every=2, final_copy=False, callback=lambda u, s: calls.append(s))
    oracle.evolve(inp)
    assert calls == [0, 2, 4]          # copy tested before steps++ (cudaOpenMP2.cu:380-386)
    assert inp.steps[0] == 5
    bad = dict(job)
    bad["omega_states"] = dict(job["omega_states"], wave_packets=job["omega_states"]["wave_packets"] * 1.01)
    with pytest.raises(RuntimeError, match="norm"):
        oracle.evolve(abi.Inputs(bad, total_steps=2))


# ---------------------------------------------------------------- flux / CRP
@pytest.mark.parametrize("n", [3, 5, 9, 21])
def test_first_derivative_weights_exact(n):
    """Stencil weights against the exact rational solution of the moment
    (Vandermonde) system sum_m c_m m^q = delta_q1, q = 0..n-1."""
    from fractions import Fraction
    half = n // 2
    offs = list(range(-half, half + 1))
    A = [[Fraction(m) ** q for m in offs] for q in range(n)]
    b = [Fraction(int(q == 1)) for q in range(n)]
    for col in range(n):        # exact Gauss-Jordan
        piv = next(r for r in range(col, n) if A[r][col] != 0)
        A[col], A[piv] = A[piv], A[col]
        b[col], b[piv] = b[piv], b[col]
        for r in range(n):
            if r != col and A[r][col] != 0:
                f = A[r][col] / A[col][col]
                A[r] = [x - f * y for x, y in zip(A[r], A[col])]
                b[r] -= f * b[col]
    exact = np.array([float(b[q] / A[q][q]) for q in range(n)])
    h = 0.05
    for got in (oracle.first_derivative(n, h), inputs.fornberg_first_derivative(n, h)):
        assert np.abs(got * h - exact).max() < 4e-16 * max(1.0, np.abs(exact).max())
    assert n != 3 or np.allclose(exact, [-0.5, 0.0, 0.5], atol=0, rtol=0)


def test_flux_eckart_transmission():
    """Known answer: the oracle's P(E) for a packet crossing a symmetric
    Eckart barrier (separable J = 0 job) equals the analytic transmission
    T(E - E_R0) to 2e-4 over T = 0.08..0.99."""
    from flux_cases import eckart_job, eckart_transmission
    job = eckart_job()
    inp = abi.Inputs(job, want_acf=False)
    oracle.evolve(inp, 8)
    T = eckart_transmission(job["meta"]["e_trans"])
    assert float(np.abs(inp.crp - T).max()) < 2e-4
    assert inp.obs[-1, 0, 0] < 1e-3      # the packet has left the grid
