"""MI355X-native split-operator wavepacket propagator (atom-diatom reactive
scattering in Jacobi coordinates with Omega-block Coriolis coupling).

The compute path is the HIP library csrc/ -> libqdamd.so behind the C ABI in
include/qd.h; this package holds the host-side pieces around it:
  inputs  - input construction (grids, LEPS potential, Legendre tables,
            initial wavepacket, absorbers, CRP normalisation)
  abi     - ctypes mirror of include/qd.h
  api     - binding to libqdamd.so (raises if the library is not built)
  serve   - partial-wave job server binding
"""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PACKAGE_DIR, "libqdamd.so")
