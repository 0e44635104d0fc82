"""Writes c4_oracle_crp.npz: the FP64 oracle's reaction probabilities for the
256 x 192 x 96, J = 0 workload with 200 collision energies in [0.01, 0.3]
hartree after 2000 steps of dt = 10 au (the flux has crossed the dividing
surface by then; 20000 steps give the same P(E)).  The mixed-precision GPU
form is checked against these values energy by energy.

The oracle (oracle/liboracle.so, built from oracle/oracle.cpp) is the FP64
restatement of the reference's step; the run takes about 15 minutes on 8
host threads.

Run: python tests/golden/make_c4_oracle.py [threads]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import conftest  # noqa: E402,F401  (loads the package and puts the repo root on sys.path)
from qd_openmp_cuda_amd import abi, inputs  # noqa: E402
from oracle import oracle  # noqa: E402

STEPS = 2000


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 8
    job = inputs.build_job(inputs.config_spec("c4", total_steps=STEPS))
    inp = abi.Inputs(job, want_acf=False, want_observables=True, final_copy=False)
    t0 = time.time()
    oracle.evolve(inp, threads)
    print(f"oracle c4 {STEPS} steps on {threads} threads: {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "c4_oracle_crp.npz"), steps=STEPS,
                        energies=job["crp"]["energies"], eta_sq=job["crp"]["eta_sq"],
                        e_int=job["meta"]["e_int"], crp=inp.crp,
                        norm_final=inp.obs[-1, :, 0].sum())


if __name__ == "__main__":
    main()
