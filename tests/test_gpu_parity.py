"""Parity of the HIP propagator (libqdamd.so, through the C ABI) with the FP64
CPU oracle on the same seeded inputs.

Bars (FP64 form): wavepacket within 1e-10 relative (L2 over all blocks) after
every checked step count, per-step observables (norm, potential, kinetic,
rotational, Coriolis energy) within 1e-9 relative to the total energy scale,
autocorrelation within 1e-10 absolute.
"""
import dataclasses
import sys

import numpy as np
import pytest

from qd_openmp_cuda_amd import abi, api, inputs
from oracle import oracle

pytestmark = pytest.mark.gpu

PSI_TOL = 1e-10
OBS_TOL = 1e-9


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _renormalised(job, target=0.99):
    """The job with its wavepacket scaled to norm `target` on the grid (coarse
    test grids under-resolve the analytic normalisation of the initial state)."""
    psi = job["omega_states"]["wave_packets"]
    w = job["theta"]["w"]
    n = float(np.sum(np.abs(psi) ** 2 * w[None, None, :, None])) * job["r1"]["dr"] * job["r2"]["dr"]
    out = dict(job)
    out["omega_states"] = dict(job["omega_states"], wave_packets=np.asfortranarray(psi * np.sqrt(target / n)))
    return out


def run_both(spec, steps, precision=abi.QD_PREC_FP64, start_steps=0, threads=8, renormalise=False):
    job = inputs.build_job(spec)
    if renormalise:
        job = _renormalised(job)
    a = abi.Inputs(job, total_steps=steps, steps=start_steps)
    b = abi.Inputs(job, total_steps=steps, steps=start_steps, precision=precision)
    _say(f"case {spec.n1}x{spec.n2}x{spec.ntheta} l_max={spec.l_max} J={spec.J} p={spec.parity} "
         f"k_max={spec.k_max} steps={steps} prec={precision}")
    oracle.evolve(a, n_threads=threads)
    api.evolve(b)
    return a, b


def rel_l2(x, y):
    return float(np.linalg.norm((x - y).ravel()) / np.linalg.norm(y.ravel()))


def check_parity(a, b, psi_tol=PSI_TOL, obs_tol=OBS_TOL):
    e_psi = rel_l2(b.psi, a.psi)
    scale = max(1.0, float(np.abs(a.obs).max()))
    e_obs = float(np.abs(b.obs - a.obs).max()) / scale
    e_acf = float(np.abs(b.acf - a.acf).max())
    _say(f"  psi rel {e_psi:.3e}  obs {e_obs:.3e}  acf {e_acf:.3e}")
    assert b.steps[0] == a.steps[0]
    assert e_psi <= psi_tol
    assert e_obs <= obs_tol
    assert e_acf <= psi_tol


def test_gemm_probe_matches_numpy():
    rng = np.random.default_rng(1)
    for rows, k, n in [(64, 4, 16), (96, 12, 20), (1024, 48, 47), (640, 63, 64), (2048, 96, 95), (100, 7, 3)]:
        A = rng.standard_normal((rows, k))
        B = rng.standard_normal((k, n))
        Cg = api.probe_gemm(A, B)
        Cr = A @ B
        err = np.abs(Cg - Cr).max() / np.abs(Cr).max()
        _say(f"gemm {rows}x{k}x{n}: {err:.2e}")
        assert err < 1e-14


@pytest.mark.parametrize("n", [24, 48, 64, 96, 128, 160, 192, 256, 384, 512])
def test_fft_probe_matches_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    f = api.probe_fft(x, inverse=False)
    g = api.probe_fft(x, inverse=True)
    ef = np.abs(f - np.fft.fft(x, axis=1)).max() / np.abs(np.fft.fft(x, axis=1)).max()
    eg = np.abs(g - n * np.fft.ifft(x, axis=1)).max() / np.abs(np.fft.fft(x, axis=1)).max()
    _say(f"fft n={n}: fwd {ef:.2e} inv {eg:.2e}")
    assert ef < 1e-14 and eg < 1e-14


def test_parity_j0_small():
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          E_coll=0.02, absorber=True, absorber_width=1.5, absorber_strength=0.1)
    a, b = run_both(spec, 6)
    check_parity(a, b)


def test_parity_coriolis_truncated():
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=5, parity=1, k_max=3, omega0=1, j0=2, E_coll=0.02)
    a, b = run_both(spec, 5)
    assert list(a.omegas) == [0, 1, 2, 3]        # J + p even: omega_min = 0, cut at k_max = 3
    check_parity(a, b)


def test_parity_coriolis_untruncated_even():
    spec = inputs.JobSpec(n1=48, n2=32, ntheta=20, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=4, parity=0, k_max=-1, omega0=0, j0=1, E_coll=0.02)
    a, b = run_both(spec, 5)
    assert list(a.omegas) == [0, 1, 2, 3, 4]
    check_parity(a, b)


def test_parity_nine_blocks():
    # K_max = 8 (the largest block count of the served workload): the 9-block R pass
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=10, parity=0, k_max=8, omega0=0, j0=2, E_coll=0.02)
    a, b = run_both(spec, 3)
    assert len(a.omegas) == 9
    check_parity(a, b)


def test_parity_nine_blocks_six_theta_tiles():
    # 96 theta points and 9 blocks: the FP64 forward Legendre pass takes the
    # paired-row kernel at 6 column tiles (multi-block jobs only), the
    # backward pass the one-row kernel
    spec = inputs.JobSpec(n1=96, n2=48, ntheta=96, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=10, parity=0, k_max=8, omega0=0, j0=2, E_coll=0.02)
    a, b = run_both(spec, 2, threads=16)
    assert len(a.omegas) == 9
    check_parity(a, b)


def test_parity_nineteen_blocks():
    # 19 blocks: the widest R-pass instantiation (up to 24 blocks in registers and LDS)
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=40, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=20, parity=0, k_max=18, omega0=0, j0=2, E_coll=0.02)
    a, b = run_both(spec, 2)
    assert len(a.omegas) == 19
    check_parity(a, b)


@pytest.mark.parametrize("waves", ["4", "8"])
def test_parity_gemm_workgroup_waves(monkeypatch, waves):
    # the Legendre GEMMs in 4- and 8-wave workgroups (large jobs take 8) on the
    # c3 block structure: 5 blocks, 32 theta points, 2 and 1 strips per wave
    monkeypatch.setenv("QD_LEG_WAVES", waves)
    spec = dataclasses.replace(inputs.config_spec("c3"), n1=96, n2=64, ntheta=32)
    a, b = run_both(spec, 3)
    check_parity(a, b)


def test_parity_theta_not_multiple_of_4():
    # n_theta = 30: both Legendre GEMMs run on k-padded operands (forward K = 30
    # in 32, backward K = 29 - Omega in 32), the forward through range-checked
    # buffer loads and the zeroed psi slabs past the last block
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=30, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=2, parity=0, k_max=-1, omega0=0, j0=1, E_coll=0.02)
    a, b = run_both(spec, 3)
    assert len(a.omegas) == 3
    check_parity(a, b)


def test_zero_steps_returns_the_inputs():
    # total_steps = 0: nothing runs (no step-0 phase either); the wavepackets
    # come back as given, time.steps stays 0 and there are no records
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=3, parity=1, k_max=-1, omega0=1, j0=1, E_coll=0.02)
    job = inputs.build_job(spec)
    _say("case zero steps 64x32x24 J=3")
    b = abi.Inputs(job, total_steps=0)
    api.evolve(b)
    assert b.steps[0] == 0
    assert np.array_equal(b.psi, job["omega_states"]["wave_packets"])
    assert b.obs.shape[0] == 0 and b.acf.shape[0] == 0


def test_parity_max_blocks_24():
    # 24 omega blocks, the most a job may have (J = 23, p = 1: omega 0..23)
    spec = inputs.JobSpec(n1=48, n2=32, ntheta=28, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=23, parity=1, k_max=-1, omega0=0, j0=2, E_coll=0.02)
    a, b = run_both(spec, 2)
    assert len(a.omegas) == 24
    check_parity(a, b)


def test_parity_runtime_R_length_512():
    # 512-point R pencils: beyond the compile-time lengths, the runtime-size R pass
    spec = inputs.JobSpec(n1=512, n2=16, ntheta=8, R_range=(2.0, 14.0), r_range=(0.5, 5.0),
                          J=1, parity=1, k_max=-1, omega0=0, j0=1, E_coll=0.02)
    a, b = run_both(spec, 2)
    check_parity(a, b)


def test_parity_long_R_many_blocks():
    # 12 blocks on 192-point R pencils: the R pass holds 2 rows of 16 blocks
    # per workgroup (the c3 path holds 4 rows of 5)
    spec = inputs.JobSpec(n1=192, n2=32, ntheta=16, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=12, parity=0, k_max=11, omega0=0, j0=2, E_coll=0.02)
    a, b = run_both(spec, 2)
    assert len(a.omegas) == 12
    check_parity(a, b)


@pytest.mark.parametrize("n2", [48, 96, 128, 192, 256])
def test_parity_r_pass_lengths(n2):
    # every compile-time length of the r pass (32 and 64 are covered above):
    # radix plans 4-3-4, 4-2-3-4, 8-2-8, 8-3-8, 8-4-8 with 16 or 8 columns per workgroup
    spec = inputs.JobSpec(n1=32, n2=n2, ntheta=8, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=1, parity=1, k_max=-1, omega0=0, j0=1, E_coll=0.02)
    a, b = run_both(spec, 3)
    check_parity(a, b)


def test_parity_runtime_fft_lengths():
    # 80 = 2^4 * 5 and 40 are not compile-time lengths: exercises the runtime-size FFT path
    spec = inputs.JobSpec(n1=80, n2=40, ntheta=20, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=2, parity=0, k_max=-1, omega0=0, j0=1, E_coll=0.02)
    a, b = run_both(spec, 4)
    check_parity(a, b)


def test_parity_c1_shape():
    spec = inputs.config_spec("c1")
    a, b = run_both(spec, 4)
    check_parity(a, b)


def test_parity_c1_long_run():
    # 600 steps of the CPU-runnable workload (128 x 64 x 48, J = 0; the packet
    # is still mostly on the grid): 8-step graphs throughout, and rounding
    # must not grow past the per-step bar over the run
    spec = inputs.config_spec("c1")
    a, b = run_both(spec, 600)
    check_parity(a, b)


def test_parity_c3_full_shape():
    # the benchmark workload itself (192 x 128 x 64, J = 20, K_max = 4, 5 blocks):
    # 8-wave GEMMs, the 192-point R passes and the 8-2-8 r pass at full size
    spec = inputs.config_spec("c3")
    a, b = run_both(spec, 2)
    assert len(a.omegas) == 5
    check_parity(a, b)


def test_parity_c3_blocks_reduced_grid():
    # the c3 Coriolis structure (J=20, K_max=4, 5 blocks) on a reduced grid
    spec = dataclasses.replace(inputs.config_spec("c3"), n1=96, n2=64, ntheta=32)
    a, b = run_both(spec, 3)
    assert len(a.omegas) == 5
    check_parity(a, b)


def test_resume_equals_single_run():
    """N steps, then M more from the stored state and step counter == N+M in one run
    (cudaOpenMP2.cu:344: V(-dt/2) only when steps == 0)."""
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=3, parity=1, k_max=-1, omega0=1, j0=1, E_coll=0.02)
    job = inputs.build_job(spec)
    one = abi.Inputs(job, total_steps=7)
    api.evolve(one)
    first = abi.Inputs(job, total_steps=3)
    api.evolve(first)
    job2 = dict(job)
    job2["omega_states"] = dict(job["omega_states"], wave_packets=first.psi.copy())
    second = abi.Inputs(job2, total_steps=4, steps=int(first.steps[0]))
    api.evolve(second)
    assert second.steps[0] == 7
    e = rel_l2(second.psi, one.psi)
    _say(f"resume: {e:.3e}")
    assert e <= 1e-13


def test_observable_ring_wraps():
    """Per-step records read back across the end of the 1024-slot device ring.
    A read at step 100 and a run on to 1100 leave records 100..1099 in slots
    100..1023 and 0..75; they equal those of one uninterrupted run, bit for bit,
    and the uninterrupted run's own norm and energies match the oracle."""
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=16, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=2, parity=0, k_max=-1, j0=1, E_coll=0.02)
    job = inputs.build_job(spec)
    _say("case ring wrap 64x32x16 J=2 steps=1100")
    one = abi.Inputs(job, total_steps=1100, final_copy=False)
    api.evolve(one)
    two = abi.Inputs(job, total_steps=1100, final_copy=False)
    with api.Propagator(two) as p:
        p.run(100)
        head = p.observables()
        p.run(1000)
        obs = p.observables()
        acf = p.autocorrelation()
    assert np.array_equal(head, one.obs[:100])
    assert np.array_equal(obs, one.obs)
    assert np.array_equal(acf, one.acf)
    ref = abi.Inputs(job, total_steps=40, final_copy=False)
    oracle.evolve(ref, 8)
    assert np.abs(one.obs[:40] - ref.obs).max() <= OBS_TOL * max(1.0, float(np.abs(ref.obs).max()))
    assert np.abs(one.acf[:40] - ref.acf).max() <= PSI_TOL


def test_fp64_bitwise_reproducible():
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, J=2, parity=0, k_max=-1, j0=1, E_coll=0.02,
                          R_range=(2.0, 12.0), r_range=(0.5, 5.0))
    job = inputs.build_job(spec)
    x = abi.Inputs(job, total_steps=4)
    y = abi.Inputs(job, total_steps=4)
    api.evolve(x)
    api.evolve(y)
    assert np.array_equal(x.psi, y.psi)
    assert np.array_equal(x.obs, y.obs)


def test_mixed_precision_close_to_fp64():
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=5, parity=1, k_max=3, omega0=1, j0=2, E_coll=0.02)
    a, b = run_both(spec, 5, precision=abi.QD_PREC_MIXED)
    e_psi = rel_l2(b.psi, a.psi)
    e_norm = float(np.abs(b.obs[..., 0].sum(1) - a.obs[..., 0].sum(1)).max())
    _say(f"  mixed: psi rel {e_psi:.3e} norm {e_norm:.3e}")
    assert e_psi < 1e-5
    assert e_norm < 1e-6


def test_copy_back_callback_matches_oracle():
    """Copy-back every 2 steps with a host callback (the reference's
    wave_to_matlab path, cudaOpenMP2.cu:380-386): the same step numbers and the
    same wavepackets as the oracle, taken before the step counter advances."""
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=3, parity=1, k_max=-1, omega0=0, j0=1, E_coll=0.02)
    job = inputs.build_job(spec)
    _say("case copy-back 64x32x24 J=3 every 2 steps, 7 steps")
    snaps = {"oracle": [], "gpu": []}

    def make(name):
        holder = {}

        def cb(user, steps):
            snaps[name].append((int(steps), holder["inp"].psi.copy()))
        inp = abi.Inputs(job, total_steps=7, copy_every=2, final_copy=False, callback=cb)
        holder["inp"] = inp
        return inp
    a, b = make("oracle"), make("gpu")
    oracle.evolve(a, n_threads=8)
    api.evolve(b)
    steps_a = [s for s, _ in snaps["oracle"]]
    steps_b = [s for s, _ in snaps["gpu"]]
    _say(f"  callback steps oracle {steps_a} gpu {steps_b}")
    assert steps_a == steps_b == [0, 2, 4, 6]
    for (_, pa), (_, pb) in zip(snaps["oracle"], snaps["gpu"]):
        assert rel_l2(pb, pa) <= PSI_TOL
    assert b.steps[0] == a.steps[0] == 7


def test_parity_without_absorber():
    # dump1/dump2 absent: the absorber is off (matlabData.h:27)
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=4, parity=1, k_max=3, omega0=1, j0=1, E_coll=0.02, absorber=False)
    a, b = run_both(spec, 5)
    assert a.d1 is None and b.d2 is None
    check_parity(a, b)


def test_command_line_driver_matches_library(tmp_path):
    """driver/qd_fixture_run (C++ over the C ABI, the counterpart of the
    reference's Matlab driver + MEX entry) gives the library's results bit for
    bit on the same fixture."""
    import os
    import subprocess
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=5, parity=1, k_max=3, omega0=1, j0=2, E_coll=0.02, flux_r=2.8,
                          n_energies=5, energy_range=(0.012, 0.03))
    job = inputs.build_job(spec)
    steps = 70
    fx, out = tmp_path / "job.bin", tmp_path / "out.bin"
    inputs.write_fixture(job, str(fx), total_steps=steps)
    drv = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qd-openmp-cuda_amd",
                       "qd_fixture_run")
    _say("case driver 64x32x24 J=5 70 steps, 5 energies (child process)")
    r = subprocess.run([drv, str(fx), str(out)], capture_output=True, text=True, timeout=120)
    _say("  " + r.stdout.strip() + r.stderr.strip())
    assert r.returncode == 0
    got = inputs.read_fixture_output(str(out), job, steps)
    ref = abi.Inputs(job, total_steps=steps, want_acf=False)
    api.evolve(ref)
    assert got["steps"] == int(ref.steps[0]) == steps
    assert np.array_equal(got["psi"], ref.psi)
    assert np.array_equal(got["obs"], ref.obs)
    assert np.array_equal(got["crp"], ref.crp)


@pytest.mark.parametrize("switch", ["QD_LEG_PAIRED", "QD_LEG_STEP", "QD_LEG_SPLIT"])
def test_legendre_kernel_forms_match_oracle(tmp_path, switch):
    """The forms of the Legendre transforms other than the default (the fused
    step with parity-split products), each selected by a switch read once per
    process, so in a child (the C++ driver): QD_LEG_PAIRED=0 the one-row-per-
    lane kernels as separate passes, QD_LEG_STEP=0 the paired-row separate
    passes, QD_LEG_SPLIT=0 the fused step with direct products.  On the c3
    block structure each matches the oracle to the same bars as the default."""
    import os
    import subprocess
    spec = dataclasses.replace(inputs.config_spec("c3"), n1=96, n2=64, ntheta=32)
    job = inputs.build_job(spec)
    steps = 3
    fx, out = tmp_path / "job.bin", tmp_path / "out.bin"
    inputs.write_fixture(job, str(fx), total_steps=steps)
    drv = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qd-openmp-cuda_amd",
                       "qd_fixture_run")
    _say(f"case {switch}=0 Legendre kernels 96x64x32 J=20 K_max=4, 3 steps (child process)")
    env = dict(os.environ, **{switch: "0"})
    r = subprocess.run([drv, str(fx), str(out)], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    got = inputs.read_fixture_output(str(out), job, steps)
    ref = abi.Inputs(job, total_steps=steps, want_acf=False)
    oracle.evolve(ref, n_threads=8)
    e_psi = rel_l2(got["psi"], ref.psi)
    e_obs = float(np.abs(got["obs"] - ref.obs).max()) / max(1.0, float(np.abs(ref.obs).max()))
    _say(f"  psi rel {e_psi:.3e}  obs {e_obs:.3e}")
    assert e_psi <= PSI_TOL and e_obs <= OBS_TOL


FFT_FORM_CASES = [
    # (grid, J, parity, k_max): the c3 block structure (5 blocks, r pass
    # with 16 columns of 64 points), 3 blocks on 128-point r pencils (8
    # columns), one block on 192-point r pencils (8 columns, two butterflies
    # a thread), 2 blocks with a 96-point R pass
    ((96, 64, 32), 20, 0, 4),
    ((64, 128, 24), 2, 0, -1),
    ((64, 192, 24), 0, 0, -1),
    ((96, 96, 24), 3, 1, -1),
    # 4 blocks in the 5-block kernels (one block absent everywhere)
    ((64, 64, 24), 4, 1, -1),
    # R pencils of 128 / 192 / 256 points: the R pass with the kinetic
    # factor on wave-local tiles (5, 3 and 1 blocks)
    ((128, 64, 24), 4, 0, -1),
    ((192, 32, 24), 2, 0, -1),
    ((256, 32, 16), 0, 0, -1),
]


@pytest.mark.parametrize("passes", [2, 3])
@pytest.mark.parametrize("case", range(len(FFT_FORM_CASES)))
def test_parity_fft_pass_forms(monkeypatch, passes, case):
    """The FFT work in two passes (the default where it has kernels: r pass
    with the first rot.C.rot half step and the r kinetic factor, R pass with
    the R kinetic factor and the second half step) and in three
    (QD_FFT_TWO_PASS=0: R pass with the first half step, r pass with both
    kinetic factors, R pass with the second half step), each against the
    oracle; the two-pass kinetic energy is summed from its R and r parts."""
    (n1, n2, nt), J, par, kmax = FFT_FORM_CASES[case]
    if passes == 3:
        monkeypatch.setenv("QD_FFT_TWO_PASS", "0")
    spec = inputs.JobSpec(n1=n1, n2=n2, ntheta=nt, R_range=(2.0, 12.0), r_range=(0.5, 5.0), J=J, parity=par,
                          k_max=kmax, omega0=par, j0=max(1, par), E_coll=0.02)
    job = inputs.build_job(spec)
    assert api.fft_passes(abi.Inputs(job, total_steps=1)) == passes
    a, b = run_both(spec, 4)
    check_parity(a, b)


def test_parity_asymmetric_theta_grid():
    """A theta grid whose nodes are not symmetric under x -> -x (one node moved
    by 1e-6; the tables as given) takes the fused step with direct products
    instead of the parity-split ones; GPU and oracle run the same inputs."""
    spec = dataclasses.replace(inputs.config_spec("c3"), n1=96, n2=64, ntheta=32)
    job = inputs.build_job(spec)
    job["theta"]["x"] = np.array(job["theta"]["x"], dtype=np.float64)
    job["theta"]["x"][0] += 1e-6
    assert api.legendre_form(abi.Inputs(job, total_steps=1)) == 1
    _say("case asymmetric theta grid 96x64x32 J=20 K_max=4, 4 steps")
    a = abi.Inputs(job, total_steps=4)
    b = abi.Inputs(job, total_steps=4)
    oracle.evolve(a, 8)
    api.evolve(b)
    check_parity(a, b)


def test_parity_c2_production_shape():
    """The 256 x 192 x 96, J = 0 production shape: the Legendre GEMMs with 6
    column tiles run the one-row-per-lane kernels in 8-wave workgroups (the
    transforms that replace the reference's ZGEMMs, omegawavepacket.cu:200-240,
    at this shape), the R passes take 256-point pencils and the r pass
    192-point pencils.  Wavepacket, observables and autocorrelation after 3
    steps against the oracle."""
    spec = inputs.config_spec("c2")
    a, b = run_both(spec, 3, threads=16)
    check_parity(a, b)


def test_parity_long_coriolis_run():
    """500 steps with Coriolis coupling: the c3 block structure (J = 20,
    K_max = 4, 5 blocks) on a 96 x 64 x 32 grid, 8-step graphs throughout.
    Rounding differences must stay under the per-step bars over the run,
    with the omega blocks mixing at every half step."""
    spec = dataclasses.replace(inputs.config_spec("c3"), n1=96, n2=64, ntheta=32)
    a, b = run_both(spec, 500, threads=16)
    assert len(a.omegas) == 5
    check_parity(a, b)
    # the packet keeps most of its norm and the blocks exchange population
    pop = a.obs[-1, :, 0]
    _say(f"  omega populations after 500 steps {np.array2string(pop, precision=4)}")
    assert pop[1:].sum() > 1e-3 * pop.sum()


@pytest.mark.parametrize("shape", [
    (16, 8, 232, 79, 0),     # forward 5 column tiles, backward 15 (the reference's n_theta = 232)
    (24, 10, 192, 79, 2),    # forward 5, backward 12, three omega blocks
    (20, 12, 160, 94, 0),    # forward 6, backward 10
])
def test_parity_wide_theta_tiles(shape):
    """n_theta above 64: the one-row-per-lane Legendre kernels at 5 to 15
    column tiles (l_max chosen so that both tables fit the LDS)."""
    n1, n2, nt, lmax, J = shape
    spec = inputs.JobSpec(n1=n1, n2=n2, ntheta=nt, l_max=lmax, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=J, parity=0, k_max=-1, omega0=0, j0=1, E_coll=0.02)
    a, b = run_both(spec, 3, renormalise=True)
    check_parity(a, b)


@pytest.mark.parametrize("waves", ["4", "8"])
@pytest.mark.parametrize("shape", [(30, 18, 24), (50, 18, 96)])
@pytest.mark.parametrize("precision", [abi.QD_PREC_FP64, abi.QD_PREC_MIXED])
def test_parity_partial_last_strip(monkeypatch, waves, shape, precision):
    """n1 n2 not a multiple of 16 (540 and 900 points): the last strip of 16
    points is partial, so the Legendre kernels clamp its loads to the last
    point and mask its stores.  FP64: the fused parity-split step at 24 theta
    points (one forward sweep) and at 96 (two sweeps); mixed: the fused step
    at 24, the one-row separate passes at 96, in 4- and 8-wave workgroups.
    Mixed against the oracle at the mixed form's bars (1e-5 relative in psi,
    1e-6 in the observables)."""
    monkeypatch.setenv("QD_LEG_WAVES", waves)
    n1, n2, nt = shape
    spec = inputs.JobSpec(n1=n1, n2=n2, ntheta=nt, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=3, parity=1, k_max=-1, omega0=1, j0=1, E_coll=0.02)
    a, b = run_both(spec, 3, precision=precision, renormalise=True)
    assert (n1 * n2) % 16 != 0
    if precision == abi.QD_PREC_FP64:
        check_parity(a, b)
    else:
        check_parity(a, b, psi_tol=1e-5, obs_tol=1e-6)


def test_autocorrelation_across_resume_matches_oracle():
    """Resumed from a stored state (steps != 0), the autocorrelation refers to
    the state at the resume time (the stored state with V(dt/2) restored):
    GPU against the oracle on the resumed run."""
    spec = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=3, parity=1, k_max=-1, omega0=1, j0=1, E_coll=0.02)
    job = inputs.build_job(spec)
    first = abi.Inputs(job, total_steps=3)
    api.evolve(first)
    job2 = dict(job)
    job2["omega_states"] = dict(job["omega_states"], wave_packets=first.psi.copy())
    _say("case resume acf 64x32x24 J=3 3+4 steps")
    a = abi.Inputs(job2, total_steps=4, steps=3)
    b = abi.Inputs(job2, total_steps=4, steps=3)
    oracle.evolve(a, 8)
    api.evolve(b)
    check_parity(a, b)


def test_mixed_long_run_norm_drift():
    """10000 steps of the J = 20, K_max = 4 workload (192 x 128 x 64, 5 blocks)
    in the mixed form beside FP64: the mixed run completes (its norm check
    never trips) and its total norm stays within 1e-6 of the FP64 norm at
    every step."""
    spec = inputs.config_spec("c3", total_steps=10000)
    job = inputs.build_job(spec)
    _say("case mixed long run 192x128x64 J=20 K_max=4 10000 steps, fp64 and mixed")
    a = abi.Inputs(job, want_acf=False, final_copy=False)
    b = abi.Inputs(job, want_acf=False, final_copy=False, precision=abi.QD_PREC_MIXED)
    api.evolve(a)
    api.evolve(b)
    na, nb = a.obs[..., 0].sum(1), b.obs[..., 0].sum(1)
    d = np.abs(nb - na)
    _say(f"  norm fp64 {na[0]:.9f} -> {na[-1]:.9f}; mixed - fp64: max {d.max():.3e} "
         f"(step {int(d.argmax())}), at 1000 {d[999]:.3e}, at 10000 {d[-1]:.3e}")
    assert b.steps[0] == 10000
    assert d.max() < 1e-6
