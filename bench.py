"""Benchmark: grid-point-steps/s of the split-operator wavepacket step on one MI355X.

Workload (the configuration the headline metric is quoted on): J = 20,
K_max = 4 (5 omega blocks), each block 192 R x 128 r x 64 theta, complex128,
LEPS surface, dt = 10 au.  A step is one full split-operator step over all
omega blocks with inputs resident in HBM.  value = total DVR grid points
(sum over blocks of n1 n2 ntheta) x steps / seconds, FP64 form; the mixed
(FP32 hi/lo Legendre) form is measured on the same workload and reported
beside it.

The serving form (qd_server_*) is measured on a seeded Poisson stream of
partial-wave jobs and reported under "serving" (grid-point-steps/s, jobs per
hour, p50/p99 latency from submit to result), FP64 and mixed.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-mixed]
                    [--no-serve] [--serve-jobs N] [--serve-steps S] [--serve-energies E]
                    [--no-configs] [--bandwidth MIB] [--c4] [--c1-batch]
"""
import argparse
import dataclasses
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "qd-openmp-cuda_amd")

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFLOPS = 78.6
# the Legendre transforms a job runs (api.legendre_form)
LEGENDRE_FORMS = {0: "separate passes", 1: "fused step, direct products", 2: "fused step, parity-split products"}
FP32_MFMA_PEAK_TFLOPS = 157.3


def _load_package():
    if "qd_openmp_cuda_amd" in sys.modules:
        return sys.modules["qd_openmp_cuda_amd"]
    spec = importlib.util.spec_from_file_location(
        "qd_openmp_cuda_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["qd_openmp_cuda_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def kernel_model(job, form=2):
    """Algorithmic bytes and flops per launch of each kernel kind for one step.

    G = n1 n2 ntheta DVR points per block, F_b = n1 n2 (l_max - omega_b + 1)
    FBR points of block b; a complex value is 16 B, the potential 8 B.  Per
    block and step: forward Legendre pass (absorber, potential phase, DGEMM)
    reads psi and V and writes the FBR slabs, 24 G + 16 F_b; each of the three
    FFT passes (R forward with the first rot.C.rot half step, r forward .
    kinetic . r inverse, R inverse with the second half step) reads and writes
    the FBR state, 32 F_b; the backward Legendre pass reads the FBR state and
    writes psi, 16 F_b + 16 G.  Total 40 G + 128 F_b.  In the two-pass form
    of the FFT work (fft_r with the first half step and the r kinetic
    factor, fft_R_inv with the R kinetic factor and the second half step)
    the two passes take 32 F_b each (step_kinds).  Flops: the real-B
    Legendre product, 4 n1 n2 ntheta N_l per block and direction.  Kinds fused
    into these kernels (dvr, fbr1, fbr2) carry no traffic of their own; the
    flux kernel runs only when reaction probabilities are requested.  The
    fused Legendre step (leg_step: the backward pass of one step with the
    forward pass of the next, psi kept on chip) reads the FBR state and V and
    writes the FBR state, 8 G + 32 F_b, with both products' flops; a step
    then moves 8 G + 128 F_b, or 8 G + 96 F_b with two FFT passes.  In its parity-split form (`form` 2, the
    default on Gauss-Legendre grids) each product runs over half the theta
    grid, so both together take 4 n1 n2 ntheta N_l per block."""
    r1, r2, th, os_ = job["r1"], job["r2"], job["theta"], job["omega_states"]
    n1, n2, nt = r1["n"], r2["n"], th["n"]
    M = n1 * n2
    G = M * nt
    omegas = [int(o) for o in os_["omegas"]]
    nO = len(omegas)
    lmax = os_["l_max"]
    F = [M * (lmax - o + 1) for o in omegas]
    Ft = sum(F)
    gemm_flops = sum(4 * M * nt * (lmax - o + 1) for o in omegas)
    zero = {"bytes": 0, "flops": 0}
    return {
        "dvr": zero,
        "leg_fwd": {"bytes": nO * 24 * G + 16 * Ft, "flops": gemm_flops},
        "fbr1": zero,
        "fft_R_fwd": {"bytes": 32 * Ft, "flops": 0},
        "fft_r": {"bytes": 32 * Ft, "flops": 0},
        "fft_R_inv": {"bytes": 32 * Ft, "flops": 0},
        "fbr2": zero,
        "leg_bwd": {"bytes": nO * 16 * G + 16 * Ft, "flops": gemm_flops},
        "flux": zero,
        "reduce": zero,
        "leg_step": {"bytes": nO * 8 * G + 32 * Ft, "flops": (1 if form == 2 else 2) * gemm_flops},
    }, G * nO


def read_traffic(kernel):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    p = os.path.join(ROOT, "profiles", "pmc_traffic_c3.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        parts = line.split()
        if len(parts) >= 2 and parts[0] == kernel:
            try:
                return float(parts[1])
            except ValueError:
                return None
    return None


def cpu_baseline(job, abi, n_steps=2):
    """FP64 oracle on the host cores over a bounded sample of the same workload
    (all threads, and one step on one thread)."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pts = job["r1"]["n"] * job["r2"]["n"] * job["theta"]["n"] * len(job["omega_states"]["omegas"])
    inp = abi.Inputs(job, total_steps=5, want_acf=False)
    oracle.evolve(inp, threads)           # 5 warm-up steps (page-in, thread pool)
    inp = abi.Inputs(job, total_steps=n_steps, want_acf=False)
    t0 = time.perf_counter()
    oracle.evolve(inp, threads)
    dt = time.perf_counter() - t0
    inp = abi.Inputs(job, total_steps=1, want_acf=False)
    t1 = time.perf_counter()
    oracle.evolve(inp, 1)
    dt1 = time.perf_counter() - t1
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    import shutil
    host = {"nproc": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0)), "model": cpu_model,
            # the reference's driver is Matlab code that needs a CUDA MEX; Octave is probed and recorded
            "octave": shutil.which("octave") or "not installed"}
    return {"value": pts * n_steps / dt, "unit": "grid-point-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n_steps} steps of the same workload after 5 warm-up steps, FP64 oracle "
                      f"(oracle/oracle.cpp, OpenMP, {threads} threads), {dt:.2f} s; {cpu_model}",
            "single_thread": {"value": pts / dt1, "sample": f"1 step on 1 thread, {dt1:.2f} s"},
            "host": host}


def mixed_crp_error(api, abi, inputs):
    """Reaction probabilities of both forms against the FP64 oracle's P(E) on
    the 256x192x96 J=0 problem at 200 energies, 2000 steps (the committed
    fixture tests/golden/c4_oracle_crp.npz, made by
    tests/golden/make_c4_oracle.py; the flux has crossed the surface by
    then).  Per energy: E_coll, eta_sq / max, the oracle's P and |dP| of
    each form; and the largest |dP| of each form over the energies where
    eta_sq is at least 1e-2, 1e-3, 1e-4 and 1e-6 of its maximum (below
    that, P = flux / eta_sq divides rounding noise by eta_sq)."""
    gold = np.load(os.path.join(ROOT, "tests", "golden", "c4_oracle_crp.npz"))
    steps = int(gold["steps"])
    job = inputs.build_job(inputs.config_spec("c4", total_steps=steps))
    assert np.array_equal(job["crp"]["energies"], gold["energies"])
    ref = gold["crp"]
    out = {}
    err = {}
    for name, prec in (("fp64", abi.QD_PREC_FP64), ("mixed", abi.QD_PREC_MIXED)):
        inp = abi.Inputs(job, precision=prec, want_acf=False, want_observables=True, final_copy=False)
        out[f"{name}_legendre_form"] = LEGENDRE_FORMS[api.legendre_form(inp)]
        t0 = time.perf_counter()
        api.evolve(inp)
        out[f"{name}_s"] = time.perf_counter() - t0
        err[name] = np.abs(inp.crp - ref)
        out[f"{name}_final_norm"] = float(inp.obs[-1, :, 0].sum())
    out["oracle_final_norm"] = float(gold["norm_final"])
    eta = job["crp"]["eta_sq"] / job["crp"]["eta_sq"].max()
    E = job["crp"]["energies"] - job["meta"]["e_int"]
    out.update({
        "workload": f"256x192x96, J=0, dt=10 au, {steps} steps, 200 collision energies in [0.01, 0.3] hartree",
        "reference": "FP64 oracle P(E), tests/golden/c4_oracle_crp.npz",
        "by_eta": [{"eta_over_max_at_least": f, "energies": int((eta >= f).sum()),
                    "max_abs_err_fp64": float(err["fp64"][eta >= f].max()),
                    "max_abs_err_mixed": float(err["mixed"][eta >= f].max())} for f in (1e-2, 1e-3, 1e-4, 1e-6)],
        "per_energy": [{"E_coll": float(e), "eta_over_max": float(h), "P_oracle": float(p),
                        "abs_err_fp64": float(d64), "abs_err_mixed": float(dmx)}
                       for e, h, p, d64, dmx in zip(E, eta, ref, err["fp64"], err["mixed"])],
    })
    return out


def step_kinds(form, passes=3):
    """Kernel kinds of one step: the fused Legendre step (forms 1, 2) or the
    separate forward and backward passes (form 0), then the FFT passes: three
    (fft_R_fwd, fft_r, fft_R_inv), or two in the two-pass form (fft_r with
    the first rot.C.rot half step and the r kinetic factor, fft_R_inv with
    the R kinetic factor and the second half step; 32 F_b each as well)."""
    leg = ("leg_step",) if form else ("leg_fwd", "leg_bwd")
    return leg + (("fft_r", "fft_R_inv") if passes == 2 else ("fft_R_fwd", "fft_r", "fft_R_inv"))


def step_model_seconds(job, form, passes=3):
    """Model time of one step: sum over its kernels of max(bytes / HBM peak,
    flops / FP64 MFMA peak), with the per-kernel figures of kernel_model."""
    model, _ = kernel_model(job, form)
    return sum(max(model[k]["bytes"] / (HBM_PEAK_GBS * 1e9), model[k]["flops"] / (FP64_MFMA_PEAK_TFLOPS * 1e12))
               for k in step_kinds(form, passes))


def other_configs(api, abi, inputs, cpu_threads, with_cpu):
    """The other two single-job configurations: c1 (128x64x48, J=0), the CPU-
    runnable case, over its full 2000 steps on the GPU and on the host (the
    oracle), and c2 (256x192x96, J=0), 1000 steps on the GPU and 50 steps on
    the host after 5 warm-up steps.  Grid-point-steps/s of each, the GPU
    step against the per-kernel HBM/MFMA model."""
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    out = {}
    for name, gpu_steps, cpu_steps, cpu_warm in (("c1", 2000, 2000, 0), ("c2", 1000, 50, 5)):
        spec = inputs.config_spec(name)
        job = inputs.build_job(spec)
        _, pts = kernel_model(job)
        rec = {"workload": f"{spec.n1}x{spec.n2}x{spec.ntheta}, J={spec.J}, {len(job['omega_states']['omegas'])} "
                           f"block(s), dt={spec.dt} au"}
        for prec_name, prec in (("fp64", abi.QD_PREC_FP64), ("mixed", abi.QD_PREC_MIXED)):
            el, _ = measure(api, abi, job, prec, gpu_steps, 1000, profile_steps=0)
            rec[prec_name] = {"steps": gpu_steps, "ms_per_step": el / gpu_steps * 1e3,
                              "grid_point_steps_per_s": pts * gpu_steps / el}
        form = api.legendre_form(abi.Inputs(job, total_steps=1))
        passes = api.fft_passes(abi.Inputs(job, total_steps=1))
        rec["fp64"]["model_frac"] = step_model_seconds(job, form, passes) / (rec["fp64"]["ms_per_step"] * 1e-3)
        rec["fp64"]["legendre_form"] = form
        rec["fp64"]["fft_passes"] = passes
        if name == "c1":
            rec["fp64_served_x16"] = c1_batched(abi, job, form, passes)
        if with_cpu:
            from oracle import oracle
            if cpu_warm:
                oracle.evolve(abi.Inputs(job, total_steps=cpu_warm, want_acf=False), cpu_threads)
            inp = abi.Inputs(job, total_steps=cpu_steps, want_acf=False, want_observables=False, final_copy=False)
            t0 = time.perf_counter()
            oracle.evolve(inp, cpu_threads)
            dt = time.perf_counter() - t0
            rec["cpu_oracle"] = {"steps": cpu_steps, "seconds": dt, "threads": cpu_threads,
                                 "grid_point_steps_per_s": pts * cpu_steps / dt,
                                 "sample": ("the full run" if cpu_steps == 2000 else
                                            f"{cpu_steps} steps after {cpu_warm} warm-up steps")}
        print(f"# {name}: fp64 {rec['fp64']['ms_per_step']:.4f} ms/step, mixed "
              f"{rec['mixed']['ms_per_step']:.4f} ms/step" +
              (f", cpu {rec['cpu_oracle']['grid_point_steps_per_s']:.3g} gps/s" if with_cpu else ""),
              file=sys.stderr, flush=True)
        out[name] = rec
    return out


def c1_batched(abi, job, form, passes, copies=16, steps=400):
    """The latency-bound 128x64x48 job batched: `copies` copies stepped
    together by the server (one launch per kernel for all of them), FP64.
    Per job-step: the wall time (admission included) over copies x steps
    job-steps, against the same per-step model as a single job's
    model_frac."""
    from qd_openmp_cuda_amd import serve
    prepared = [(0.0, abi.Inputs(job, total_steps=steps, want_acf=False, final_copy=False, calculate_crp=False))
                for _ in range(copies)]
    r = serve.run_stream(prepared, time_scale=0.0)
    per_job_step = r["wall_s"] / (copies * steps)
    return {"jobs": copies, "steps": steps, "wall_s": r["wall_s"],
            "grid_point_steps_per_s": r["grid_point_steps_per_s"], "ms_per_job_step": per_job_step * 1e3,
            "model_frac": step_model_seconds(job, form, passes) / per_job_step}


def served_crp_error(crp64, crpmx, etas):
    """Reaction probabilities of the served mixed form against the served FP64
    form, job by job (FP64 served results equal qd_evolve alone bitwise and
    match the oracle in the tests): max |dP| over the energies where the job's
    eta_sq is at least 1e-2 / 1e-3 of its maximum."""
    out = {}
    for frac in (1e-2, 1e-3):
        worst, n = 0.0, 0
        for a, b, e in zip(crp64, crpmx, etas):
            ok = e >= frac * e.max()
            if ok.any():
                worst = max(worst, float(np.abs(a[ok] - b[ok]).max()))
                n += int(ok.sum())
        out[f"eta_ge_{frac:.0e}_max"] = {"energies": n, "max_abs_err": worst}
    return out


def measure(api, abi, job, precision, steps, warmup, profile_steps=20):
    inp = abi.Inputs(job, total_steps=steps + warmup, precision=precision, want_acf=False,
                     want_observables=False, final_copy=False)
    with api.Propagator(inp) as p:
        p.run(warmup)
        p.synchronize()
        t0 = time.perf_counter()
        p.run(steps)
        p.synchronize()
        elapsed = time.perf_counter() - t0
        prof = p.profile(profile_steps) if profile_steps else None
    return elapsed, prof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    # the warm-up also brings the GPU to its working clocks (measured: c3 at
    # 0.234 ms/step after 20 warm-up steps, 0.199 after 2000)
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=50)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the c1 / c2 GPU and CPU figures")
    ap.add_argument("--serve-energies", type=int, default=50,
                    help="reaction probabilities per served job (0: none)")
    ap.add_argument("--serve-only", default="", choices=("", "fp64_saturated", "mixed_saturated", "fp64_poisson",
                                                         "mixed_poisson"),
                    help="only run this serving case and print its figures (profiling)")
    ap.add_argument("--workload", default="c3", choices=("c1", "c2", "c3"),
                    help="single-job workload of the timed steps (profiling; the headline is c3)")
    ap.add_argument("--shape", default="", help="profiling: N1xN2xNTHETA grid in place of the workload's")
    ap.add_argument("--blocks", type=int, default=0,
                    help="profiling: omega blocks (J = 2 (blocks - 1), K_max = blocks - 1) with --shape")
    ap.add_argument("--bandwidth", type=int, default=0,
                    help="only run the HBM ceiling probes over this many MiB and exit")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--c1-batch", action="store_true",
                    help="only run 16 copies of the 128x64x48 job through the server (profiling) and exit")
    ap.add_argument("--c4", action="store_true",
                    help="only run the mixed-precision CRP comparison (256x192x96, J=0, 200 energies) and exit")
    ap.add_argument("--serve-jobs", type=int, default=64)
    ap.add_argument("--serve-steps", type=int, default=200)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain copy of this many MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus N > 1 is not supported. Partial-wave jobs are "
                                   "independent, so a multi-GPU deployment runs one server per GPU and "
                                   "deals jobs out by memory footprint"}))
        sys.exit(2)
    _load_package()
    from qd_openmp_cuda_amd import abi, api, inputs
    if args.copy_probe:
        nbytes = args.copy_probe << 20
        us = api.probe_copy(nbytes, 10)
        print(json.dumps({"copy_bytes": nbytes, "us_per_copy": us, "GB/s": 2 * nbytes / us / 1e3}))
        return
    if args.bandwidth:
        nbytes = args.bandwidth << 20
        res = {"bytes": nbytes}
        us = api.probe_copy(nbytes, 10)
        res["plain_copy"] = {"us": us, "GB/s": 2 * nbytes / us / 1e3}
        us = api.probe_bandwidth(nbytes, 1, 10)
        res["wide_copy_nt"] = {"us": us, "GB/s": 2 * nbytes / us / 1e3}
        us = api.probe_bandwidth(nbytes, 2, 10)
        res["read_only"] = {"us": us, "GB/s": nbytes / us / 1e3}
        print(json.dumps(res))
        return

    if args.c4:
        print(json.dumps(mixed_crp_error(api, abi, inputs)))
        return
    if args.c1_batch:
        job1 = inputs.build_job(inputs.config_spec("c1"))
        inp1 = abi.Inputs(job1, total_steps=1)
        print(json.dumps(c1_batched(abi, job1, api.legendre_form(inp1), api.fft_passes(inp1))))
        return

    if args.serve_only:
        from qd_openmp_cuda_amd import serve
        name, mode = args.serve_only.split("_")
        stream = inputs.serve_jobs(n_jobs=args.serve_jobs, total_steps=args.serve_steps,
                                   n_energies=args.serve_energies)
        prepared = serve.prepare(stream, precision=abi.QD_PREC_MIXED if name == "mixed" else abi.QD_PREC_FP64,
                                 want_crp=args.serve_energies > 0)
        r = serve.run_stream(prepared, time_scale=0.0 if mode == "saturated" else 1.0)
        r.pop("latency_s")
        print(json.dumps(r))
        return

    spec = inputs.config_spec(args.workload)
    if args.shape:
        n1, n2, nt = (int(v) for v in args.shape.lower().split("x"))
        nb = max(args.blocks, 1)
        spec = dataclasses.replace(spec, n1=n1, n2=n2, ntheta=nt, J=2 * (nb - 1), parity=0, k_max=nb - 1,
                                   R0=9.0, seed=None)
    job = inputs.build_job(spec)
    form = api.legendre_form(abi.Inputs(job, total_steps=1))
    passes = api.fft_passes(abi.Inputs(job, total_steps=1))
    model, pts = kernel_model(job, form)
    print(f"# workload {args.workload}: {spec.n1}x{spec.n2}x{spec.ntheta}, J={spec.J}, K_max={spec.k_max}, "
          f"{len(job['omega_states']['omegas'])} blocks; steps={args.steps} warmup={args.warmup}",
          file=sys.stderr, flush=True)

    el64, prof64 = measure(api, abi, job, abi.QD_PREC_FP64, args.steps, args.warmup)
    value = pts * args.steps / el64
    print(f"# fp64: {el64 / args.steps * 1e3:.3f} ms/step, {value:.4g} gps/s", file=sys.stderr, flush=True)

    mixed = None
    if not args.no_mixed:
        elm, profm = measure(api, abi, job, abi.QD_PREC_MIXED, args.steps, args.warmup)
        mixed = {"value": pts * args.steps / elm, "ms_per_step": elm / args.steps * 1e3,
                 "ratio_vs_fp64": el64 / elm, "kernels_us": {k: round(v, 2) for k, v in profm.items()}}
        print(f"# mixed: {elm / args.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)

    # roofline of the dominant kernel (largest time per step); its bound is the
    # larger of bytes / HBM peak and flops / MFMA peak for that kernel
    dom = max(prof64, key=lambda k: prof64[k])
    m = model[dom]
    us = prof64[dom]
    if m["flops"] / (FP64_MFMA_PEAK_TFLOPS * 1e12) > m["bytes"] / (HBM_PEAK_GBS * 1e9):
        achieved = m["flops"] / (us * 1e-6) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / FP64_MFMA_PEAK_TFLOPS}
    else:
        achieved = m["bytes"] / (us * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS}
    roof["traffic"] = read_traffic(dom)
    roof["kernel"] = dom
    roof["legendre_form"] = LEGENDRE_FORMS[form]
    roof["fft_passes"] = passes
    roof["kernel_us"] = us
    roof["algorithmic_bytes"] = m["bytes"]
    roof["algorithmic_flops"] = m["flops"]
    # whole step against the same per-kernel model: sum of max(bytes/BW, flops/peak)
    t_model = step_model_seconds(job, form, passes)
    roof["step_model_us"] = t_model * 1e6
    roof["step_model_frac"] = t_model / (el64 / args.steps)
    # the same step against the five-pass model (separate Legendre passes,
    # 40G + 128F per block): the bytes a step moves without the fused step
    t_5pass = step_model_seconds(job, 0)
    roof["step_model_5pass_us"] = t_5pass * 1e6
    roof["step_model_5pass_frac"] = t_5pass / (el64 / args.steps)
    per_kernel = {}
    for k, us_k in prof64.items():
        if us_k <= 0.0:         # kinds fused into a neighbouring kernel
            continue
        mk = model[k]
        per_kernel[k] = {"us": round(us_k, 2),
                         "GB/s": round(mk["bytes"] / (us_k * 1e-6) / 1e9, 1) if us_k > 0 else None,
                         "TFLOP/s": round(mk["flops"] / (us_k * 1e-6) / 1e12, 2) if mk["flops"] else None}

    serving = None
    if not args.no_serve:
        serving = {}
        from qd_openmp_cuda_amd import serve
        stream = inputs.serve_jobs(n_jobs=args.serve_jobs, total_steps=args.serve_steps,
                                   n_energies=args.serve_energies)
        want_crp = args.serve_energies > 0
        crp = {}
        # "poisson": the seeded arrival times (throughput bounded by the arrival
        # rate); "saturated": every request at t = 0 (the server's capacity)
        for name, prec in (("fp64", abi.QD_PREC_FP64), ("mixed", abi.QD_PREC_MIXED)):
            for mode, scale in (("poisson", 1.0), ("saturated", 0.0)):
                prepared = serve.prepare(stream, precision=prec, want_crp=want_crp)
                r = serve.run_stream(prepared, time_scale=scale)
                if mode == "saturated" and want_crp:
                    crp[name] = [inp.crp.copy() for _, inp in prepared]
                    crp["eta"] = [inp.eta.copy() for _, inp in prepared]
                del prepared
                r.pop("latency_s")
                serving[f"{name}_{mode}"] = r
                print(f"# serving {name} {mode}: {r['grid_point_steps_per_s']:.4g} gps/s, "
                      f"p50 {r['latency_p50_s']:.2f} s, p99 {r['latency_p99_s']:.2f} s, {r['failed']} failed",
                      file=sys.stderr, flush=True)
        if want_crp:
            serving["mixed_vs_fp64_crp"] = served_crp_error(crp["fp64"], crp["mixed"], crp["eta"])
        # the same 64 jobs one at a time (qd_run, device-resident, FP64): the
        # time the saturated server would need at the single-job rates
        alone = 0.0
        for _, spec in stream:
            alone += measure(api, abi, inputs.build_job(spec), abi.QD_PREC_FP64, spec.total_steps, 0,
                             profile_steps=0)[0]
        serving["fp64_alone_s"] = alone
        serving["fp64_saturated_vs_alone"] = alone / serving["fp64_saturated"]["wall_s"]
        print(f"# serving fp64: the jobs one at a time take {alone:.2f} s of steps; saturated server at "
              f"{serving['fp64_saturated_vs_alone']:.2f} of that rate", file=sys.stderr, flush=True)
        serving["workload"] = (f"{args.serve_jobs} Poisson arrivals (4/s, seed 5), J uniform 0..40, "
                               f"K_max=min(J,8), shapes 128x64x48 / 192x128x64 / 256x192x96, "
                               f"{args.serve_steps} steps each, reaction probabilities at "
                               f"{args.serve_energies} energies per job")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(job, abi, args.cpu_steps)

    configs = None
    if not args.no_configs:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        configs = other_configs(api, abi, inputs, threads, not args.no_cpu)

    out = {
        "metric": "grid-point-steps/s, FP64 vs mixed, J=20 K=4; p50/p99 partial-wave latency",
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el64 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": ("J=20, K_max=4, 5 omega blocks of 192x128x64 (R x r x theta), complex128, "
                                "LEPS surface, dt=10 au") if args.workload == "c3" else
                               f"{args.workload}: {spec.n1}x{spec.n2}x{spec.ntheta}, J={spec.J}, complex128",
                   "grid_points_per_step": pts, "parallelism": "single GPU, all omega blocks resident"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "serving": serving,
        "kernels": per_kernel,
        "configs": configs,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
