"""ctypes mirror of include/qd.h and a builder that fills `qd_inputs` from
the arrays produced by inputs.build_job.

The struct layouts here must match include/qd.h field for field; the
field order follows the reference's input struct (FH2main.m:195-206,
matlabStructures.h:12-167)."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

QD_OK, QD_ERR_INVALID, QD_ERR_DEVICE, QD_ERR_NORM, QD_ERR_STATE, QD_ERR_NOMEM, QD_ERR_CAPACITY = range(7)
QD_PREC_FP64, QD_PREC_MIXED = 0, 1
QD_JOB_QUEUED, QD_JOB_RUNNING, QD_JOB_DONE, QD_JOB_FAILED = range(4)
QD_OBS_FIELDS = 5
OBS_NAMES = ("norm", "potential", "kinetic", "rotational", "coriolis")
QD_KERNEL_KINDS = 11
# kernel kinds of qd_profile_steps (dvr, fbr1 and fbr2 are fused into
# leg_fwd / leg_step, fft_R_fwd and fft_R_inv and read 0; flux runs only with
# CRP; where the fused Legendre step runs, leg_fwd and leg_bwd read 0 and the
# reduction runs inside fft_R_fwd, so reduce reads 0 as well)
KERNEL_NAMES = ("dvr", "leg_fwd", "fbr1", "fft_R_fwd", "fft_r", "fft_R_inv", "fbr2", "leg_bwd",
                "flux", "reduce", "leg_step")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)

WAVE_CALLBACK = C.CFUNCTYPE(None, C.c_void_p, C.c_int32)


class qd_radial(C.Structure):
    _fields_ = [("n", C.c_int32), ("left", C.c_double), ("dr", C.c_double), ("mass", C.c_double)]


class qd_angle(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("x", _dp), ("w", _dp),
                ("associated_legendre", _dp)]


class qd_time(C.Structure):
    _fields_ = [("total_steps", C.c_int32), ("time_step", C.c_double), ("steps", _ip)]


class qd_options(C.Structure):
    _fields_ = [("wave_to_matlab", C.c_char_p), ("test_name", C.c_char_p),
                ("steps_to_copy_psi_from_device_to_host", C.c_int32),
                ("use_p2p_async", C.c_int32),
                ("wave_callback", WAVE_CALLBACK), ("callback_user", C.c_void_p)]


class qd_dump(C.Structure):
    _fields_ = [("dump", _dp)]


class qd_crp(C.Structure):
    _fields_ = [("energies", _dp), ("eta_sq", _dp), ("CRP", _dp),
                ("n_dividing_surface", C.c_int32), ("n_gradient_points", C.c_int32),
                ("n_energies", C.c_int32), ("calculate_CRP", C.c_int32)]


class qd_omega_states(C.Structure):
    _fields_ = [("J", C.c_int32), ("parity", C.c_int32), ("l_max", C.c_int32),
                ("n_omegas_max", C.c_int32), ("n_omegas", C.c_int32), ("omegas", _ip),
                ("legendre_n_max", C.c_int32), ("associated_legendres", _dp),
                ("wave_packets", _dp)]


class qd_inputs(C.Structure):
    _fields_ = [("r1", qd_radial), ("r2", qd_radial), ("theta", qd_angle), ("pot", _dp),
                ("time", qd_time), ("options", qd_options), ("dump1", qd_dump),
                ("dump2", qd_dump), ("crp", qd_crp), ("omega_states", qd_omega_states),
                ("k_max", C.c_int32), ("precision", C.c_int32), ("print_energies", C.c_int32),
                ("final_copy", C.c_int32), ("observables", _dp), ("autocorrelation", _dp)]


def _ptr(a: Optional[np.ndarray], ct=_dp):
    if a is None:
        return C.cast(None, ct)
    return a.ctypes.data_as(ct)


class Inputs:
    """Owns the numpy buffers behind one `qd_inputs` struct.

    `job` is a dict from inputs.build_job.  The wavepacket buffer is a private
    copy (in/out), exposed as `self.psi` (complex, (n1, n2, ntheta, n_omegas),
    Fortran order).  Output buffers for observables / autocorrelation / CRP
    are allocated for `total_steps` steps."""

    def __init__(self, job: dict, total_steps: Optional[int] = None, steps: Optional[int] = None,
                 precision: int = QD_PREC_FP64, want_observables: bool = True,
                 want_acf: bool = True, final_copy: bool = True, copy_every: int = 0,
                 callback=None, print_energies: bool = False, calculate_crp: Optional[bool] = None):
        self.job = job
        r1, r2, th, os_ = job["r1"], job["r2"], job["theta"], job["omega_states"]
        self.x = np.ascontiguousarray(th["x"], dtype=np.float64)
        self.w = np.ascontiguousarray(th["w"], dtype=np.float64)
        self.pot = np.asfortranarray(job["pot"], dtype=np.float64)
        self.P = np.asfortranarray(os_["associated_legendres"], dtype=np.float64)
        self.omegas = np.ascontiguousarray(os_["omegas"], dtype=np.int32)
        self.psi = np.array(os_["wave_packets"], dtype=np.complex128, order="F", copy=True)
        self.d1 = None if job["dump1"] is None else np.ascontiguousarray(job["dump1"], dtype=np.float64)
        self.d2 = None if job["dump2"] is None else np.ascontiguousarray(job["dump2"], dtype=np.float64)
        ts = job["time"]["total_steps"] if total_steps is None else total_steps
        self.steps = np.array([job["time"]["steps"] if steps is None else steps], dtype=np.int32)
        nO = len(self.omegas)
        self.obs = np.zeros((ts, nO, QD_OBS_FIELDS)) if want_observables else None
        self.acf = np.zeros(ts, dtype=np.complex128) if want_acf else None
        crp = job["crp"]
        self.energies = np.ascontiguousarray(crp["energies"], dtype=np.float64)
        self.eta = np.ascontiguousarray(crp["eta_sq"], dtype=np.float64)
        self.crp = np.zeros(len(self.energies))
        do_crp = crp["calculate_CRP"] if calculate_crp is None else int(bool(calculate_crp))

        s = qd_inputs()
        s.r1 = qd_radial(r1["n"], r1["left"], r1["dr"], r1["mass"])
        s.r2 = qd_radial(r2["n"], r2["left"], r2["dr"], r2["mass"])
        s.theta = qd_angle(th["n"], th["m"], _ptr(self.x), _ptr(self.w), _ptr(None))
        s.pot = _ptr(self.pot)
        s.time = qd_time(ts, job["time"]["time_step"], self.steps.ctypes.data_as(_ip))
        self._cb = WAVE_CALLBACK(callback) if callback is not None else WAVE_CALLBACK()
        s.options = qd_options(None, None, copy_every, 0, self._cb, None)
        s.dump1 = qd_dump(_ptr(self.d1))
        s.dump2 = qd_dump(_ptr(self.d2))
        s.crp = qd_crp(_ptr(self.energies), _ptr(self.eta), _ptr(self.crp), crp["n_dividing_surface"],
                       crp["n_gradient_points"], len(self.energies), do_crp)
        s.omega_states = qd_omega_states(os_["J"], os_["parity"], os_["l_max"], nO, nO,
                                         self.omegas.ctypes.data_as(_ip), self.P.shape[1],
                                         _ptr(self.P), self.psi.ctypes.data_as(_dp))
        s.k_max = job.get("k_max", -1)
        s.precision = precision
        s.print_energies = int(print_energies)
        s.final_copy = int(final_copy)
        s.observables = _ptr(self.obs)
        s.autocorrelation = C.cast(self.acf.ctypes.data, _dp) if self.acf is not None else _ptr(None)
        self.struct = s

    def ref(self):
        return C.byref(self.struct)
