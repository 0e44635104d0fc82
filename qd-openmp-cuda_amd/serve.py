"""Request loop of the serving form.

Feeds a timed arrival stream of partial-wave jobs to the device server
(include/qd.h, qd_server_*): requests are submitted when their arrival time
comes, while earlier jobs are being stepped on the device (continuous
batching); the server admits them as memory frees up and retires each job
as soon as it has run its steps.  Reports per-job latency (submit to result)
and the aggregate grid-point-steps/s.

All device work goes through libqdamd; this loop only keeps time.
"""
from __future__ import annotations

import time
from typing import List, Sequence, Tuple

import numpy as np

from . import abi, api, inputs


def job_points(inp: abi.Inputs) -> int:
    """DVR grid points per step of one job (sum over its omega blocks)."""
    s = inp.struct
    return s.r1.n * s.r2.n * s.theta.n * s.omega_states.n_omegas


def prepare(stream: Sequence[Tuple[float, "inputs.JobSpec"]], precision: int = abi.QD_PREC_FP64,
            want_crp: bool = False) -> List[Tuple[float, abi.Inputs]]:
    """Build the inputs of every request ahead of the clock (the client's work)."""
    out = []
    for t, spec in stream:
        job = inputs.build_job(spec)
        out.append((t, abi.Inputs(job, precision=precision, want_acf=False, want_observables=True,
                                  final_copy=True, calculate_crp=want_crp)))
    return out


def run_stream(prepared: Sequence[Tuple[float, abi.Inputs]], budget_bytes: int = 0, max_running: int = 0,
               time_scale: float = 1.0, max_steps: int = 8) -> dict:
    """Replay the arrivals (seconds, scaled by time_scale) against one server.

    Each server call advances every shape class by up to max_steps steps
    (qd_server_advance).  Returns per-job latency and queueing times,
    p50/p99, and the aggregate throughput over the span from the first
    submission to the last result."""
    order = sorted(range(len(prepared)), key=lambda q: prepared[q][0])
    with api.Server(budget_bytes, max_running) as srv:
        ids = {}
        active = set()
        nxt = 0
        t0 = time.perf_counter()
        steps = 0
        while nxt < len(order) or active:
            now = time.perf_counter() - t0
            while nxt < len(order) and prepared[order[nxt]][0] * time_scale <= now:
                q = order[nxt]
                jid = srv.submit(prepared[q][1])
                ids[q] = jid
                active.add(q)
                nxt += 1
            if active:
                srv.step(max_steps)
                steps += 1
                for q in list(active):
                    if srv.status(ids[q])["state"] in (api.Server.DONE, api.Server.FAILED):
                        active.discard(q)
            elif nxt < len(order):
                time.sleep(max(0.0, min(0.01, prepared[order[nxt]][0] * time_scale - now)))
        wall = time.perf_counter() - t0
        stats = [srv.status(ids[q]) for q in range(len(prepared))]
        errors = {q: srv.error(ids[q]) for q in range(len(prepared)) if stats[q]["state"] == api.Server.FAILED}
    lat = np.array([s["latency_s"] for s in stats])
    qd = np.array([s["queued_s"] for s in stats])
    pts = sum(job_points(inp) * inp.struct.time.total_steps for _, inp in prepared)
    return {
        "jobs": len(prepared),
        "failed": len(errors),
        "errors": errors,
        "wall_s": wall,
        "server_steps": steps,
        "grid_point_steps": pts,
        "grid_point_steps_per_s": pts / wall,
        "jobs_per_hour": len(prepared) / wall * 3600.0,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "queued_p50_s": float(np.percentile(qd, 50)),
        "latency_s": lat.tolist(),
    }
