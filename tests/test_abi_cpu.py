"""C-ABI checks that need no GPU: the library loads, exports every symbol
declared in include/*.h, rejects invalid inputs before touching the device
(the reference's insist() conditions, the odd-n case of cudaMath.h:119,
omegawavepacket.cu:46/:142, CoriolisMatrix.C:65, cudaOpenMP2.cu:437-440),
and its host-side Coriolis eigensolver agrees with the oracle's."""
import ctypes as C
import dataclasses
import os
import re

import numpy as np
import pytest

from qd_openmp_cuda_amd import abi, api, inputs
from oracle import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(qd_[a-z0-9_]+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = api.lib()
    names = declared_functions()
    assert "qd_evolve" in names and "qd_create" in names and "qd_probe_gemm" in names
    for n in names:
        assert hasattr(L, n), n
    assert b"gfx950" in L.qd_version()


def test_struct_layout_matches_header():
    # field offsets as a C compiler lays out include/qd.h
    assert C.sizeof(abi.qd_radial) == 32
    assert abi.qd_inputs.r2.offset == 32
    assert abi.qd_inputs.theta.offset == 64


def small_job(**kw):
    base = dict(n1=32, n2=24, ntheta=12, R_range=(2.0, 10.0), r_range=(0.6, 4.0), E_coll=0.01)
    base.update(kw)
    return inputs.build_job(inputs.JobSpec(**base))


def expect_invalid(inp, pattern):
    with pytest.raises(api.QDError) as ei:
        api.evolve(inp)
    assert ei.value.code == abi.QD_ERR_INVALID, str(ei.value)
    assert re.search(pattern, str(ei.value)), str(ei.value)


def test_rejects_odd_grid():
    job = small_job()
    inp = abi.Inputs(job, total_steps=1)
    inp.struct.r1.n = 31
    expect_invalid(inp, "even")


def test_rejects_theta_not_exceeding_lmax():
    job = small_job()
    inp = abi.Inputs(job, total_steps=1)
    inp.struct.omega_states.l_max = 11
    expect_invalid(inp, "l_max")


def test_rejects_nonconsecutive_omegas():
    job = small_job(J=3, parity=1, k_max=-1, omega0=1, j0=1)
    inp = abi.Inputs(job, total_steps=1)
    inp.omegas[1] = 5
    expect_invalid(inp, "consecutive")


def test_rejects_parity_for_j0():
    job = small_job()
    inp = abi.Inputs(job, total_steps=1)
    inp.struct.omega_states.parity = 1
    expect_invalid(inp, "parity")


def test_rejects_incomplete_omega_list():
    job = small_job(J=3, parity=1, k_max=2, omega0=1, j0=1)      # omegas 1..2
    inp = abi.Inputs(job, total_steps=1)
    inp.struct.k_max = -1                                         # claims untruncated J=3
    expect_invalid(inp, "min\\(J, l_max, k_max\\)")


def test_rejects_norm_above_one():
    job = small_job()
    job["omega_states"] = dict(job["omega_states"], wave_packets=job["omega_states"]["wave_packets"] * 1.001)
    expect_invalid(abi.Inputs(job, total_steps=1), "norm")


def test_rejects_bad_flux_surface():
    job = small_job(n_energies=4, flux_r=0.7)
    expect_invalid(abi.Inputs(job, total_steps=1), "stencil")


def test_rejects_missing_copy_interval():
    job = small_job()
    inp = abi.Inputs(job, total_steps=1, callback=lambda u, s: None, copy_every=0)
    expect_invalid(inp, "steps_to_copy")


def test_null_inputs():
    L = api.lib()
    assert L.qd_evolve(None) == abi.QD_ERR_INVALID
    h = C.c_void_p()
    assert L.qd_create(None, C.byref(h)) == abi.QD_ERR_INVALID
    assert L.qd_run(None, 1) == abi.QD_ERR_INVALID


@pytest.mark.parametrize("J,l,om,n", [(1, 1, 0, 2), (20, 30, 0, 5), (8, 40, 1, 8), (40, 50, 1, 8), (23, 23, 0, 24)])
def test_host_eigensolver_matches_oracle(J, l, om, n):
    e, v = api.probe_coriolis(J, l, om, n)
    c = oracle.coriolis_dense(J, l, om, n)
    eo, _ = oracle.sym_eig(c)
    scale = max(1.0, np.abs(c).max())
    assert np.allclose(np.sort(e), np.sort(eo), atol=1e-11 * scale)
    assert np.allclose(v @ np.diag(e) @ v.T, c, atol=1e-11 * scale)
    assert np.allclose(v.T @ v, np.eye(n), atol=1e-13)


DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qd-openmp-cuda_amd",
                      "qd_fixture_run")


def test_driver_usage_and_bad_fixture(tmp_path):
    """The command-line driver links the library and rejects what is not a
    fixture before touching the device."""
    import subprocess
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"not a fixture at all")
    r = subprocess.run([DRIVER, str(bad), str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "not a fixture" in r.stderr


def test_fixture_header_roundtrip(tmp_path):
    spec = inputs.JobSpec(n1=16, n2=16, ntheta=6, J=2, parity=0, k_max=-1, j0=1, n_energies=3,
                          r_range=(0.5, 5.0), flux_r=2.0, n_gradient_points=5)
    job = inputs.build_job(spec)
    p = tmp_path / "f.bin"
    inputs.write_fixture(job, str(p), total_steps=5)
    raw = p.read_bytes()
    assert raw[:8] == b"QDFX0001"
    h = np.frombuffer(raw[8:8 + 17 * 4], dtype=np.int32)
    assert list(h[:5]) == [16, 16, 6, 4, 3] and h[9] == 5 and h[12] == 3 and h[16] == 1
    n1, n2, nt, nO, lnmax = 16, 16, 6, 3, int(h[8])
    expect = 8 + 17 * 4 + 7 * 8 + 8 * (2 * nt + n1 * n2 * nt) + 4 * nO + 8 * nt * lnmax * nO \
        + 16 * n1 * n2 * nt * nO + 8 * (n1 + n2) + 8 * 2 * 3
    assert len(raw) == expect


def test_legendre_form_choice():
    """The Legendre transforms a job runs (host-side choice, no GPU): the fused
    step with parity-split products where the theta grid and the tables are
    symmetric under x -> -x (Gauss-Legendre nodes, normalised associated
    Legendre functions), the direct fused step where they are not, the
    separate passes where the fused kernels have no instantiation.  The
    mixed form has the split fused step only (on an asymmetric grid it runs
    the separate passes), and not at 3 theta tiles per half (96 points)."""
    job = inputs.build_job(inputs.config_spec("c3"))
    assert api.legendre_form(abi.Inputs(job, total_steps=1)) == 2

    def altered(x=None, P=None, precision=abi.QD_PREC_FP64):
        j = dict(job)
        j["theta"] = dict(job["theta"], x=job["theta"]["x"] if x is None else x)
        if P is not None:
            j["omega_states"] = dict(job["omega_states"], associated_legendres=P)
        return abi.Inputs(j, total_steps=1, precision=precision)

    # one node moved by 1e-6: not symmetric, direct form
    x = np.array(job["theta"]["x"], dtype=np.float64)
    x[0] += 1e-6
    assert api.legendre_form(altered(x=x)) == 1
    # one table entry off its mirror image by 1e-9 of the row's largest value
    P = np.array(job["omega_states"]["associated_legendres"], dtype=np.float64, order="F")
    P[3, 5, 1] += 1e-9 * np.abs(P[:, 5, 1]).max()
    assert api.legendre_form(altered(P=P)) == 1
    # a difference at rounding level keeps the split form
    P = np.array(job["omega_states"]["associated_legendres"], dtype=np.float64, order="F")
    P[3, 5, 1] *= 1.0 + 1e-15
    assert api.legendre_form(altered(P=P)) == 2
    assert api.legendre_form(abi.Inputs(job, total_steps=1, precision=abi.QD_PREC_MIXED)) == 2
    assert api.legendre_form(altered(x=x, precision=abi.QD_PREC_MIXED)) == 0
    big = inputs.build_job(inputs.config_spec("c2"))
    assert api.legendre_form(abi.Inputs(big, total_steps=1)) == 2
    assert api.legendre_form(abi.Inputs(big, total_steps=1, precision=abi.QD_PREC_MIXED)) == 0
    # 128 theta points (4 tiles per half): no instantiation of either fused form
    wide = inputs.build_job(inputs.JobSpec(n1=32, n2=16, ntheta=128, R_range=(2.0, 12.0), r_range=(0.5, 5.0)))
    assert api.legendre_form(abi.Inputs(wide, total_steps=1)) == 0


def test_fft_pass_choice(monkeypatch):
    """The FFT work's pass count (host-side choice, no GPU): two passes where
    both lengths have compile-time kernels and the r pass with the first half
    step has a tiling for the blocks (up to 5), three otherwise or with
    QD_FFT_TWO_PASS=0."""
    def passes(spec):
        return api.fft_passes(abi.Inputs(inputs.build_job(spec), total_steps=1))
    assert passes(inputs.config_spec("c3")) == 2          # 192 x 128, 5 blocks
    assert passes(inputs.config_spec("c2")) == 2          # 256 x 192, 1 block
    assert passes(inputs.config_spec("c1")) == 2          # 128 x 64, 1 block
    nine = inputs.JobSpec(n1=64, n2=32, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=8, parity=0, k_max=8, omega0=0, j0=1)
    assert passes(nine) == 3                              # 9 blocks
    wide = inputs.JobSpec(n1=64, n2=192, ntheta=24, R_range=(2.0, 12.0), r_range=(0.5, 5.0),
                          J=4, parity=0, k_max=-1, omega0=0, j0=1)
    assert passes(wide) == 3                              # 5 blocks of 192-point r pencils: no tiling
    odd = inputs.JobSpec(n1=40, n2=32, ntheta=12, R_range=(2.0, 12.0), r_range=(0.5, 5.0))
    assert passes(odd) == 3                               # R length without compile-time kernels
    monkeypatch.setenv("QD_FFT_TWO_PASS", "0")
    assert passes(inputs.config_spec("c3")) == 3
