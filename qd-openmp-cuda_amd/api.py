"""Binding to libqdamd.so (the HIP propagator behind include/qd.h).

The library is the only compute path: if it is not built, `lib()` raises.
There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import LIB_PATH
from . import abi

_lib = None

_DECLS = {
    "qd_evolve": (C.c_int, [C.c_void_p]),
    "qd_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "qd_run": (C.c_int, [C.c_void_p, C.c_int32]),
    "qd_synchronize": (C.c_int, [C.c_void_p]),
    "qd_copy_wavepackets": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qd_get_observables": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "qd_get_autocorrelation": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "qd_get_crp": (C.c_int, [C.c_void_p, C.c_void_p]),
    "qd_steps_done": (C.c_int32, [C.c_void_p]),
    "qd_destroy": (None, [C.c_void_p]),
    "qd_last_error": (C.c_char_p, [C.c_void_p]),
    "qd_last_global_error": (C.c_char_p, []),
    "qd_profile_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "qd_version": (C.c_char_p, []),
    "qd_probe_gemm": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
    "qd_probe_fft": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "qd_fft_pencils_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "qd_probe_coriolis": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    "qd_probe_copy": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "qd_probe_bandwidth": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_double)]),
    "qd_probe_legendre_form": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "qd_probe_fft_passes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "qd_server_create": (C.c_int, [C.c_int64, C.c_int32, C.POINTER(C.c_void_p)]),
    "qd_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]),
    "qd_server_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "qd_server_advance": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "qd_server_job_status": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "qd_server_job_error": (C.c_char_p, [C.c_void_p, C.c_int32]),
    "qd_server_memory": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "qd_server_destroy": (None, [C.c_void_p]),
    "qd_server_last_error": (C.c_char_p, [C.c_void_p]),
}


class QDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"qd error {code}: {msg}")
        self.code = code


def lib(path: Optional[str] = None):
    """Load libqdamd.so; raises FileNotFoundError if it has not been built."""
    global _lib
    if _lib is None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise FileNotFoundError(f"{p} is not built (run __graft_entry__.build())")
        _lib = C.CDLL(p)
        for name, (res, args) in _DECLS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


def _check(rc: int, h=None):
    if rc != abi.QD_OK:
        L = lib()
        msg = (L.qd_last_error(h) if h else L.qd_last_global_error()) or b""
        raise QDError(rc, msg.decode())


def evolve(inputs: abi.Inputs) -> None:
    """qd_evolve: the one-shot drop-in (create, run total_steps, copy back, destroy)."""
    _check(lib().qd_evolve(C.addressof(inputs.struct)))


class Propagator:
    """A resident job (qd_create / qd_run / ... / qd_destroy)."""

    def __init__(self, inputs: abi.Inputs):
        self.inputs = inputs
        h = C.c_void_p()
        _check(lib().qd_create(C.addressof(inputs.struct), C.byref(h)))
        self.h = h

    def run(self, n_steps: int) -> None:
        _check(lib().qd_run(self.h, n_steps), self.h)

    def synchronize(self) -> None:
        _check(lib().qd_synchronize(self.h), self.h)

    @property
    def steps_done(self) -> int:
        return lib().qd_steps_done(self.h)

    def wavepackets(self) -> np.ndarray:
        psi = np.zeros_like(self.inputs.psi)
        _check(lib().qd_copy_wavepackets(self.h, psi.ctypes.data), self.h)
        return psi

    def observables(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        n = self.steps_done - first if count is None else count
        nO = len(self.inputs.omegas)
        out = np.zeros((n, nO, abi.QD_OBS_FIELDS))
        _check(lib().qd_get_observables(self.h, first, n, out.ctypes.data), self.h)
        return out

    def autocorrelation(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        n = self.steps_done - first if count is None else count
        out = np.zeros(n, dtype=np.complex128)
        _check(lib().qd_get_autocorrelation(self.h, first, n, out.ctypes.data), self.h)
        return out

    def crp(self) -> np.ndarray:
        out = np.zeros(len(self.inputs.energies))
        _check(lib().qd_get_crp(self.h, out.ctypes.data), self.h)
        return out

    def profile(self, n_steps: int) -> dict:
        us = np.zeros(abi.QD_KERNEL_KINDS)
        _check(lib().qd_profile_steps(self.h, n_steps, us.ctypes.data), self.h)
        return dict(zip(abi.KERNEL_NAMES, us.tolist()))

    def close(self) -> None:
        if self.h:
            lib().qd_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving form (qd_server_*): many resident jobs stepped together.

    submit() queues an abi.Inputs (kept alive here until the job retires);
    step() admits what fits the memory budget, advances every running job by
    one step and retires finished ones (their outputs land in their Inputs)."""

    QUEUED, RUNNING, DONE, FAILED = abi.QD_JOB_QUEUED, abi.QD_JOB_RUNNING, abi.QD_JOB_DONE, abi.QD_JOB_FAILED

    def __init__(self, budget_bytes: int = 0, max_running: int = 0):
        h = C.c_void_p()
        rc = lib().qd_server_create(budget_bytes, max_running, C.byref(h))
        if rc != abi.QD_OK:
            raise QDError(rc, (lib().qd_server_last_error(None) or b"").decode())
        self.h = h
        self._inputs = {}

    def _check(self, rc):
        if rc != abi.QD_OK:
            raise QDError(rc, (lib().qd_server_last_error(self.h) or b"").decode())

    def submit(self, inputs: abi.Inputs) -> int:
        jid = C.c_int32()
        self._check(lib().qd_server_submit(self.h, C.addressof(inputs.struct), C.byref(jid)))
        self._inputs[jid.value] = inputs
        return jid.value

    def step(self, max_steps: int = 1):
        """qd_server_step (max_steps = 1) or qd_server_advance."""
        nr, nq = C.c_int32(), C.c_int32()
        if max_steps == 1:
            self._check(lib().qd_server_step(self.h, C.byref(nr), C.byref(nq)))
        else:
            self._check(lib().qd_server_advance(self.h, max_steps, C.byref(nr), C.byref(nq)))
        return nr.value, nq.value

    def status(self, job_id: int):
        st, done, lat, q = C.c_int32(), C.c_int32(), C.c_double(), C.c_double()
        self._check(lib().qd_server_job_status(self.h, job_id, C.byref(st), C.byref(done), C.byref(lat),
                                               C.byref(q)))
        return {"state": st.value, "steps": done.value, "latency_s": lat.value, "queued_s": q.value}

    def error(self, job_id: int) -> str:
        return (lib().qd_server_job_error(self.h, job_id) or b"").decode()

    def memory(self):
        u, b = C.c_int64(), C.c_int64()
        self._check(lib().qd_server_memory(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def close(self):
        if self.h:
            lib().qd_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# probes (include/qd_test.h)
# ---------------------------------------------------------------------------

def legendre_form(inp) -> int:
    """Form of the Legendre transforms a job would run (host only): 0 the
    separate passes, 1 the fused step, 2 the fused step with parity-split
    products."""
    form = C.c_int32(-1)
    _check(lib().qd_probe_legendre_form(C.addressof(inp.struct), C.byref(form)))
    return form.value


def fft_passes(inp) -> int:
    """Read-write passes over the FBR array of a job's FFT work (host only):
    3, or 2 in the two-pass form (r pass with the first half step, R pass
    with the R kinetic factor)."""
    n = C.c_int32(-1)
    _check(lib().qd_probe_fft_passes(C.addressof(inp.struct), C.byref(n)))
    return n.value


def probe_gemm(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    rows, k = A.shape
    n = B.shape[1]
    Cm = np.zeros((rows, n), order="F")
    _check(lib().qd_probe_gemm(A.ctypes.data, rows, k, B.ctypes.data, n, Cm.ctypes.data))
    return Cm


def probe_fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """x: (count, n) complex; returns the unnormalised DFT of each row."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    count, n = a.shape
    _check(lib().qd_probe_fft(a.ctypes.data, n, count, int(inverse)))
    return a


def fft_pencils_noalloc(data_ptr: int, n: int, count: int, inverse: bool = False, stream: int = 0) -> None:
    """qd_fft_pencils_noalloc: in-place DFT of `count` pencils of length n at
    the DEVICE address data_ptr, enqueued on `stream` (0: the default stream)."""
    _check(lib().qd_fft_pencils_noalloc(C.c_void_p(data_ptr), n, count, int(inverse), C.c_void_p(stream or None)))


def probe_coriolis(J: int, l: int, omega_min: int, n: int):
    e = np.zeros(n)
    v = np.zeros((n, n), order="F")
    _check(lib().qd_probe_coriolis(J, l, omega_min, n, e.ctypes.data, v.ctypes.data))
    return e, v


def probe_bandwidth(nbytes: int, kind: int, repeats: int = 10) -> float:
    """Microseconds per pass of the HBM ceiling probes: kind 1 wide copy of
    nbytes (nontemporal stores), kind 2 read-only pass over nbytes."""
    us = C.c_double()
    _check(lib().qd_probe_bandwidth(nbytes, repeats, kind, C.byref(us)))
    return us.value


def probe_copy(nbytes: int, repeats: int = 10) -> float:
    """Microseconds per plain device copy of nbytes (read nbytes + write nbytes)."""
    us = C.c_double()
    _check(lib().qd_probe_copy(nbytes, repeats, C.byref(us)))
    return us.value
