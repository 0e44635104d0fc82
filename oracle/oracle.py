"""ctypes loader for the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> None:
    args = ["make", "-C", HERE] + (["-B"] if force else [])
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.oracle_evolve.restype = C.c_int
        _lib.oracle_evolve.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
        _lib.oracle_fft1d.argtypes = [C.c_int, C.c_int, C.c_void_p]
        _lib.oracle_sym_eig.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.oracle_coriolis_dense.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        _lib.oracle_coriolis_exp.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                             C.c_void_p]
        _lib.oracle_first_derivative.argtypes = [C.c_int, C.c_double, C.c_void_p]
    return _lib


def evolve(inputs, n_threads: int = 0) -> int:
    """Run the oracle on an abi.Inputs (same contract as qd_evolve).
    Raises RuntimeError with the oracle's message on failure."""
    err = C.create_string_buffer(256)
    rc = lib().oracle_evolve(C.addressof(inputs.struct), n_threads, err)
    if rc != 0:
        raise RuntimeError(f"oracle_evolve failed ({rc}): {err.value.decode()}")
    return rc


def fft1d(a, sign: int):
    import numpy as np
    x = np.array(a, dtype=np.complex128, copy=True)
    lib().oracle_fft1d(x.size, sign, x.ctypes.data)
    return x


def sym_eig(a):
    import numpy as np
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    e = np.zeros(n)
    v = np.zeros((n, n), order="F")
    lib().oracle_sym_eig(n, a.ctypes.data, e.ctypes.data, v.ctypes.data)
    return e, v   # v[:, alpha] is eigenvector alpha


def coriolis_dense(J: int, l: int, omega_min: int, n: int):
    import numpy as np
    c = np.zeros((n, n), order="F")
    lib().oracle_coriolis_dense(J, l, omega_min, n, c.ctypes.data)
    return c


def coriolis_exp(J: int, l: int, omega_min: int, n: int, tau: float, I: float):
    import numpy as np
    out = np.zeros((n, n), dtype=np.complex128)
    lib().oracle_coriolis_exp(J, l, omega_min, n, tau, I, out.ctypes.data)
    return out


def first_derivative(n_points: int, h: float):
    import numpy as np
    c = np.zeros(n_points)
    lib().oracle_first_derivative(n_points, h, c.ctypes.data)
    return c
