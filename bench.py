"""Throughput of the TM5 transport step on one MI355X: simulated years per hour at glb3x2x34.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step here is one meteo interval of 3600 s on the glb3x2 global grid (120 x 90 x 34, one CH4
tracer, FP64): Check_CFL (the mass-only rehearsal that may cut ndyn), then its n steps of ndyn,
each Setup_MassFlow + advection x/y/z with sub-cycling, convection with Kzz diffusion, CH4
injection and chemistry, m2phlb1 -- the state resident in HBM, inputs synthetic (seeded,
pydyn_amd.meteo).  Every interval is bit-identical to the reference (tests/test_gpu_parity.py).

value = simulated seconds advanced / wall seconds x 3600 / (365 x 86400), the reference's
years/hour (pydyn/gpu/bench.py:26, 262) with the ndyn Check_CFL actually chose.

Also reported: the dominant kernel's roofline (its algorithmic bytes per launch over its
average launch time, measured here with HIP events on the stream it runs on) and the oracle
(NumPy, one core) timed on a bounded sample of the same workload on this host.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

SECONDS_A_YEAR = 365.0 * 86400.0
METRIC = "simulated years/hour at glb3x2x34 (FP64 bit-exact); 64-member member-steps/s; HBM GB/s"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=48, help="timed intervals of 3600 s")
    p.add_argument("--warmup", type=int, default=4)
    p.add_argument("--grid", default="glb3x2")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-intervals", type=int, default=4)
    return p.parse_args()


def time_kernel(fn, reps, torch):
    """Average duration (s) of fn() over reps launches, with events on the current stream."""
    fn()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(reps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / 1e3 / reps


def roofline(reg, torch):
    """The x advection (dynamu) at glb3x2 on copies of the region's state: algorithmic bytes
    per launch = read and write of m, rm, rxm, rym, rzm (im jm lm each) and am ((im+1) jm lm),
    divided by the average launch time."""
    from pydyn_amd import device as D
    from pydyn_amd.gpu import kernels as K

    g = reg.g
    m, am = reg.m.clone(), reg.am.clone()
    q = [t.clone() for t in (reg.rm, reg.rxm, reg.rym, reg.rzm)]
    saved = [t.clone() for t in (m, am, *q)]

    def once():
        K.advect_x(m, *q, am, deferred=True)

    def restore():
        for dst, src in zip((m, am, *q), saved):
            dst.copy_(src)

    cells = g.im * g.jm * g.lm
    faces = (g.im + 1) * g.jm * g.lm
    nt = q[0].shape[3]
    traffic = 2 * 8 * (cells * (1 + 4 * nt) + faces)
    # time single launches with their state restored in between (the restore is not timed)
    durations = []
    for _ in range(20):
        restore()
        durations.append(time_kernel(once, 1, torch))
    D.context().raise_errors()
    dt = float(np.median(durations))
    achieved = traffic / dt / 1e9
    return {
        "kernel": "k_x (dynamu, x advection with sub-cycling)",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None,
        "bytes_per_launch": traffic,
        "launch_us": round(dt * 1e6, 2),
    }


def cpu_baseline(met, intervals):
    from oracle.driver import Run

    run = Run(met)
    t = time.perf_counter()
    sim = 0
    for _ in range(intervals):
        ndyn, n = run.interval()
        sim += ndyn * n
    wall = time.perf_counter() - t
    return {
        "value": round(sim / wall * 3600.0 / SECONDS_A_YEAR, 6),
        "unit": "simulated years/hour",
        "cores": 1,
        "kind": "port",
        "sample": f"{intervals} intervals ({sim} simulated s) of glb3x2x34 from t=0, the NumPy oracle, "
        f"single-threaded (os.cpu_count()={os.cpu_count()}, affinity {len(os.sched_getaffinity(0))})",
        "wall_s": round(wall, 2),
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (members/tracers would shard "
              "across GPUs with no transport-time communication)", file=sys.stderr)
        return 2
    import torch

    from pydyn_amd import meteo
    from pydyn_amd.region import Region

    met = meteo.Meteo(meteo.grid(a.grid), seed=1)
    reg = Region(met)
    for _ in range(a.warmup):
        reg.interval()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim = 0
    for _ in range(a.steps):
        ndyn, n = reg.interval()
        sim += ndyn * n
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    reg.check()
    value = sim / wall * 3600.0 / SECONDS_A_YEAR
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "simulated years/hour",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded pydyn_amd.meteo: steady balanced circulation, CFL 0.3-1.5)",
        "config": {
            "workload": f"{a.grid} x {met.g.lm} plain FP64 interval (Check_CFL + steps), 1 CH4 tracer",
            "grid": [met.g.im, met.g.jm, met.g.lm],
            "lmax": met.g.lmax,
            "nt": met.nt,
            "interval_s": 3600,
            "simulated_s": sim,
        },
        "roofline": roofline(reg, torch),
    }
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(met, a.cpu_intervals)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
