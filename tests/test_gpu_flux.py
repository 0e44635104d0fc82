"""Reaction probabilities from the device flux path (surface kernel, windowed
energy transform, finish) against the FP64 oracle and a known answer.

Bars: P(E) within 1e-9 relative of the oracle at every energy, plus an
absolute floor of 1e-15 of the largest |P| for energies where P passes
through zero (FP64 form); the Eckart case also within 2e-4 of the analytic
transmission.
"""
import dataclasses
import sys

import numpy as np
import pytest

from qd_openmp_cuda_amd import abi, api, inputs
from oracle import oracle
from flux_cases import eckart_job, eckart_transmission

pytestmark = pytest.mark.gpu

CRP_TOL = 1e-9


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def crp_rel_err(gpu, ref, floor=1e-6):
    """max_e |P_gpu - P_ref| / (|P_ref| + floor max|P_ref|).  Against a 1e-9
    bar the floor allows 1e-15 max|P| of rounding noise where P passes through
    zero (P is the imaginary part of a sum whose terms are far larger); a 1%
    error at any energy with |P| above 1e-13 of the largest fails."""
    scale = float(np.abs(ref).max())
    return float((np.abs(gpu - ref) / (np.abs(ref) + floor * scale)).max())


def _multiblock_job(n_energies=10):
    # J = 5, p = 1, K_max = 3: four omega blocks; surface r = 2.8 bohr keeps the
    # 21-point stencil inside the 32-point r grid
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=5, parity=1, k_max=3, omega0=1, j0=2, E_coll=0.02, flux_r=2.8,
                          n_energies=n_energies, energy_range=(0.012, 0.03))
    return inputs.build_job(spec)


def test_flux_eckart_matches_oracle_and_transmission():
    job = eckart_job()
    _say("case eckart 32x512x4 J=0 steps=1600 energies=12")
    a = abi.Inputs(job, want_acf=False)
    b = abi.Inputs(job, want_acf=False)
    oracle.evolve(a, 8)
    api.evolve(b)
    T = eckart_transmission(job["meta"]["e_trans"])
    e_or = float(np.abs(b.crp - a.crp).max())
    e_an = float(np.abs(b.crp - T).max())
    _say(f"  crp vs oracle {e_or:.3e}  vs analytic {e_an:.3e}")
    assert e_or < CRP_TOL
    assert e_an < 2e-4


def test_flux_multiblock_parity():
    """Four omega blocks, 100 steps: one full 64-step window plus a partial one."""
    job = _multiblock_job()
    _say("case flux 64x32x24 J=5 p=1 k_max=3 steps=100 energies=10")
    a = abi.Inputs(job, total_steps=100)
    b = abi.Inputs(job, total_steps=100)
    oracle.evolve(a, 8)
    api.evolve(b)
    # after 100 steps the flux has not converged and P is small (about 4e-9):
    # the bar is relative per energy
    e = crp_rel_err(b.crp, a.crp)
    _say(f"  crp vs oracle {e:.3e} relative  (max |P| {np.abs(a.crp).max():.3e}, "
         f"max |dP| {np.abs(b.crp - a.crp).max():.3e})")
    assert np.abs(a.crp).max() > 0.0
    assert e < CRP_TOL
    assert np.abs(b.psi - a.psi).max() < 1e-10 * np.abs(a.psi).max()


def test_multi_step_graphs_match_single_launches(monkeypatch):
    """Runs of 100 and then 37 more steps replay 8-step graphs between the
    host-side events (64-step flux window flushes, the end of a run); the
    wavepackets, observables and P(E) equal a run that launches every step
    on its own (no graphs) bit for bit."""
    job = _multiblock_job()
    _say("case flux 64x32x24 J=5 p=1 k_max=3 steps=100+37 graphs vs plain launches")
    res = []
    for no_graph in ("0", "1"):
        monkeypatch.setenv("QD_NO_GRAPH", no_graph)
        b = abi.Inputs(job, total_steps=137)
        with api.Propagator(b) as p:
            p.run(100)
            p.run(37)
            res.append((p.wavepackets(), p.observables(), p.crp(), int(b.steps[0])))
    (pa, oa, ca, sa), (pb, ob, cb, sb) = res
    _say(f"  max |dpsi| {np.abs(pa - pb).max():.3e}  max |dobs| {np.abs(oa - ob).max():.3e}  "
         f"max |dP| {np.abs(ca - cb).max():.3e}")
    assert sa == sb
    assert np.array_equal(pa, pb)
    assert np.array_equal(oa, ob)
    assert np.array_equal(ca, cb)


def test_crp_readout_mid_run():
    """qd_get_crp folds the open window; reading it after 50 steps and again
    after 100 gives the 100-step one-shot result."""
    job = _multiblock_job(n_energies=6)
    _say("case flux mid-run readout 64x32x24 steps=50+50")
    one = abi.Inputs(job, total_steps=100)
    api.evolve(one)
    inp = abi.Inputs(job, total_steps=100)
    with api.Propagator(inp) as p:
        p.run(50)
        half = p.crp()
        p.run(50)
        full = p.crp()
    ref = abi.Inputs(job, total_steps=50)
    api.evolve(ref)
    _say(f"  50: {np.abs(half - ref.crp).max():.3e}  100: {np.abs(full - one.crp).max():.3e}")
    assert np.abs(half - ref.crp).max() <= 1e-12 * max(1.0, np.abs(ref.crp).max())
    assert np.abs(full - one.crp).max() <= 1e-12 * max(1.0, np.abs(one.crp).max())


def test_crp_not_requested_is_an_error():
    job = _multiblock_job(n_energies=0)
    inp = abi.Inputs(job, total_steps=2)
    with api.Propagator(inp) as p:
        p.run(2)
        with pytest.raises(api.QDError):
            lib_out = np.zeros(1)
            rc = api.lib().qd_get_crp(p.h, lib_out.ctypes.data)
            api._check(rc, p.h)


def _resolved(job, frac):
    eta = job["crp"]["eta_sq"]
    return eta >= frac * eta.max()


def test_flux_c2_shape_parity_200_energies():
    """The 256x192x96 J=0 production shape with 200 energies, 20 steps: P(E)
    against the oracle, relative per energy, wherever the initial packet has
    energy content; the wavepacket and the observables as in the parity tests."""
    job = inputs.build_job(inputs.config_spec("c4", total_steps=20))
    _say("case flux 256x192x96 J=0 steps=20 energies=200")
    a = abi.Inputs(job, want_acf=False)
    b = abi.Inputs(job, want_acf=False)
    oracle.evolve(a, 16)
    api.evolve(b)
    ok = _resolved(job, 1e-8)
    # after 20 steps only the tail of the packet reaches the surface: P is the
    # transform of an exponentially growing signal whose sums cancel to 1e-7
    # of their terms, and the device (64-step windows as products) and the
    # oracle (step by step) sum in different orders, so the bar here is 1e-8
    # relative; the converged P(E) of this shape is held to 1e-9 against the
    # oracle's 2000-step fixture in test_mixed_crp_error_c4_shape
    e = crp_rel_err(b.crp[ok], a.crp[ok], floor=1e-3)
    e_psi = float(np.linalg.norm(b.psi - a.psi) / np.linalg.norm(a.psi))
    e_obs = float(np.abs(b.obs - a.obs).max()) / max(1.0, float(np.abs(a.obs).max()))
    _say(f"  crp vs oracle {e:.3e} relative over {int(ok.sum())} energies (max |P| "
         f"{np.abs(a.crp[ok]).max():.3e})  psi rel {e_psi:.3e}  obs {e_obs:.3e}")
    assert e < 10 * CRP_TOL
    assert e_psi < 1e-10
    assert e_obs < 1e-9


def test_mixed_crp_error_c4_shape():
    """The 256x192x96 J=0 problem, 2000 steps (the flux has passed the surface
    by then), 200 energies, against the FP64 oracle's P(E) (the committed
    fixture tests/golden/c4_oracle_crp.npz, written by make_c4_oracle.py):
      FP64 form: within 1e-9 relative wherever eta_sq >= 1e-4 of its maximum
        (below that P = flux / eta_sq divides rounding noise by eta_sq);
      mixed form (FP32 hi/lo Legendre products): |dP| <= 1e-6 wherever
        eta_sq >= 1e-2 of its maximum; the error at every energy is printed,
        it grows like 1/eta_sq outside the packet's energy range."""
    import os
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c4_oracle_crp.npz"))
    steps = int(gold["steps"])
    job = inputs.build_job(inputs.config_spec("c4", total_steps=steps))
    assert np.array_equal(job["crp"]["energies"], gold["energies"])
    _say(f"case mixed crp 256x192x96 J=0 steps={steps} energies=200 against the oracle fixture")
    a = abi.Inputs(job, want_acf=False, want_observables=False, final_copy=False)
    b = abi.Inputs(job, want_acf=False, want_observables=False, final_copy=False, precision=abi.QD_PREC_MIXED)
    api.evolve(a)
    api.evolve(b)
    ref = gold["crp"]
    eta = job["crp"]["eta_sq"] / job["crp"]["eta_sq"].max()
    e64 = np.abs(a.crp - ref)
    emx = np.abs(b.crp - ref)
    for q in range(0, 200, 5):
        _say(f"  E_coll {job['crp']['energies'][q] - job['meta']['e_int']:.4f}  eta/max {eta[q]:.1e}  "
             f"P {ref[q]:+.6e}  |dP| fp64 {e64[q]:.2e} mixed {emx[q]:.2e}")
    for frac in (1e-2, 1e-3, 1e-4, 1e-6):
        ok = eta >= frac
        _say(f"  eta >= {frac:.0e} max ({int(ok.sum())} energies): max |dP| fp64 {e64[ok].max():.3e}  "
             f"mixed {emx[ok].max():.3e}")
    ok64 = eta >= 1e-4
    assert crp_rel_err(a.crp[ok64], ref[ok64]) < CRP_TOL
    assert emx[eta >= 1e-2].max() <= 1e-6


def test_reaction_probability_within_bounds_at_dt_5():
    """0 <= P(E) <= 1 for the 128x64x48 J=0 F + H2 job propagated with dt = 5 au
    over 20000 au (4000 steps; the norm left on the grid is 2e-5), wherever the
    initial packet carries energy (eta_sq >= 1e-3 of its maximum).  At dt = 10 au
    the split-operator step error alone moves P(E) by up to 0.1 and past 1 at
    some energies, so the bound is checked at the smaller step."""
    spec = inputs.config_spec("c1", dt=5.0, total_steps=4000, n_energies=24, energy_range=(0.02, 0.09))
    job = inputs.build_job(spec)
    _say("case flux bounds 128x64x48 J=0 dt=5 steps=4000 energies=24")
    b = abi.Inputs(job, want_acf=False, want_observables=False, final_copy=False)
    api.evolve(b)
    ok = _resolved(job, 1e-3)
    _say(f"  P over {int(ok.sum())} energies: min {b.crp[ok].min():.4f} max {b.crp[ok].max():.4f}")
    assert b.crp[ok].min() >= 0.0
    assert b.crp[ok].max() <= 1.0
