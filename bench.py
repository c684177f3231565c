"""Throughput of the TM5 transport step on one MI355X: simulated years per hour at glb3x2x34.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step here is one meteo interval of 3600 s on the glb3x2 global grid (120 x 90 x 34, one CH4
tracer, FP64): Check_CFL (the mass-only rehearsal that may cut ndyn), then its n steps of ndyn,
each Setup_MassFlow + advection x/y/z with sub-cycling, convection with Kzz diffusion, CH4
injection and chemistry, m2phlb1 -- the state resident in HBM, inputs synthetic (seeded,
pydyn_amd.meteo).  Every interval is bit-identical to the reference (tests/test_gpu_parity.py).

value = simulated seconds advanced / wall seconds x 3600 / (365 x 86400), the reference's
years/hour (pydyn/gpu/bench.py:26, 262) with the ndyn Check_CFL actually chose.

Also reported, in the same line: the zoom tree's rate ("zoom"), the 64-member x 4-tracer batched
form on it against one member's plain run ("batched"), the cost of one gradient over 24 hourly
intervals through the adjoint ("adjoint"), the dominant kernel's roofline (its algorithmic bytes or FP64 operations per
launch over its average launch time, from HIP events the library records around each launch on
its stream, over intervals run after the timed ones) and the oracle (NumPy, one core) timed on a
bounded sample of the same workload on this host.
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
    p.add_argument("--cpu-intervals", type=int, default=24)
    p.add_argument("--profile-intervals", type=int, default=12, help="intervals timed per kernel after the run")
    p.add_argument("--no-zoom", action="store_true", help="skip the zoom tree's rate")
    p.add_argument("--no-batched", action="store_true", help="skip the 64-member batched rate")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint window's cost")
    p.add_argument("--members", type=int, default=64)
    p.add_argument("--only", choices=("zoom", "batched", "adjoint"),
                   help="run that section alone and print its object (for profiling one workload)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="end with three plain device copies of MIB MiB (a known size for rocprofv3 --pmc)")
    return p.parse_args()


FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)


def column_flops(n, nt):
    """Algorithmic FP64 operations of the fused column kernel per column: the dense Crout LU
    of the n x n convection matrix (n-1-k divisions and (n-1-k)^2 multiply-subtracts at step
    k), its explicit inverse by forward and back substitution of the n unit vectors, and per
    tracer two applications (rm <- C rm, three slopes times diag C).  The matrix build is not
    counted."""
    lu = sum(m + 2 * m * m for m in range(n))
    inverse = n * (2 * n * (n - 1) + n)
    apply = nt * 2 * (2 * n * n + 3 * n)
    return lu + inverse + apply


def kernel_units(g, nt):
    """Algorithmic work per launch of the step's kernels at grid g: (bound, amount, unit).
    Bytes are one read of every input and one write of every output array the kernel touches
    (sub-cycles stay on chip)."""
    cells = g.im * g.jm * g.lm
    fx, fy, fz = (g.im + 1) * g.jm * g.lm, g.im * (g.jm + 1) * g.lm, g.im * g.jm * (g.lm + 1)
    cols = g.im * g.jm
    q = 4 * nt
    return {
        "k_column<%d>" % g.lmax: ("mfma", cols * column_flops(g.lmax, nt), "flop"),
        "k_x<false>": ("hbm", 8 * (2 * cells * (1 + q) + 2 * fx), "B"),
        "k_y_steps<false>": ("hbm", 8 * (2 * cells * (1 + q) + fy), "B"),
        "k_z_steps<false>": ("hbm", 8 * (2 * cells * (1 + q) + fz), "B"),
        "k_x<true>": ("hbm", 8 * (2 * cells + 2 * fx), "B"),
        "k_y_steps<true>": ("hbm", 8 * (2 * cells + fy), "B"),
        "k_z_steps<true>": ("hbm", 8 * (2 * cells + fz), "B"),
        "k_z_search": ("hbm", 8 * (cells + fz), "B"),
        "k_y_search": ("hbm", 8 * (cells + fy), "B"),
        "k_dynam0": ("hbm", 8 * (fx + fy + fx + fy + fz), "B"),
    }


def traffic_of(kernel):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy kernel of known size), as recorded in profiles/traffic.json; None if not recorded."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return None if rec is None else rec.get("bytes_per_launch")


def roofline(reg, intervals):
    """Per-kernel GPU time of `intervals` more intervals, from events the library records around
    every launch on its stream; the roofline of the kernel with the largest total."""
    from pydyn_amd import device as D

    if intervals < 1:
        return None
    ctx = D.context()
    ctx.profile_begin()
    for _ in range(intervals):
        reg.interval()
    prof = ctx.profile_end()
    total = sum(s for _, s in prof.values())
    kernel, (count, secs) = max(prof.items(), key=lambda kv: kv[1][1])
    units = kernel_units(reg.g, reg.rm.shape[3])
    if kernel not in units:
        return {"kernel": kernel, "bound": None, "achieved": None, "peak": None, "unit": None, "frac": None,
                "traffic": None, "launch_us": round(secs / count * 1e6, 2)}
    bound, amount, unit = units[kernel]
    avg = secs / count
    if bound == "hbm":
        achieved, peak, runit = amount / avg / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        achieved, peak, runit = amount / avg / 1e12, FP64_PEAK_TFLOPS, "TFLOP/s"
    out = {
        "kernel": kernel,
        "bound": bound,
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": runit,
        "frac": round(achieved / peak, 4),
        "traffic": traffic_of(kernel),
        "algorithmic_per_launch": amount,
        "algorithmic_unit": unit,
        "launch_us": round(avg * 1e6, 2),
        "launches": count,
        "share_of_gpu_time": round(secs / total, 3),
    }
    if bound == "mfma":
        out["note"] = ("FP64 vector adds and multiplies, no FMA (bit-exactness); peak is the published FP64 "
                       "figure, which counts an FMA as two operations")
    return out


def _timed(run, intervals, warmup=2):
    """Wall seconds of `intervals` intervals after `warmup`, and the parent steps and simulated
    seconds they advanced."""
    import torch

    for _ in range(warmup):
        run.interval()
    torch.cuda.synchronize()
    t = time.perf_counter()
    steps = sim = 0
    for _ in range(intervals):
        ndyn, n = run.interval()
        steps += n
        sim += ndyn * n
    torch.cuda.synchronize()
    return time.perf_counter() - t, steps, sim


def zoom_section(intervals=12):
    """The zoom tree (glb3x2x34 parent + the 1x1 degree child over 170W-50W, 10N-80N,
    120 x 70 x 34, tref 2), one CH4 tracer: simulated years/hour."""
    from pydyn_amd import meteo
    from pydyn_amd.zoom.run import ZoomRun

    p, c = meteo.zoom("c2")
    run = ZoomRun(p, c)
    wall, steps, sim = _timed(run, intervals)
    return {
        "workload": "glb3x2x34 + child 120x70x34 (xref 3, yref 2, tref 2), two-way exchange, 1 CH4 tracer",
        "value": round(sim / wall * 3600.0 / SECONDS_A_YEAR, 3),
        "unit": "simulated years/hour",
        "ms_per_interval": round(wall / intervals * 1e3, 3),
        "intervals": intervals,
    }


def batched_section(members, intervals=6):
    """members x 4 tracers on the zoom tree in one run (each member's emission scaled, each
    bit-identical to its own plain run) against one member's plain run: member-steps/s, a step
    being one parent step of the tree."""
    import torch

    from pydyn_amd import meteo
    from pydyn_amd.zoom.run import ZoomRun

    p, c = meteo.zoom("c2", nt=4)
    ens = meteo.Ensemble(p, members)
    plain = ZoomRun(p, c, ensemble=ens.subset([0]))
    wall_p, steps_p, _ = _timed(plain, 2 * intervals)
    del plain
    torch.cuda.empty_cache()
    batched = ZoomRun(p, c, ensemble=ens)
    wall_b, steps_b, _ = _timed(batched, intervals)
    sequential = steps_p / wall_p  # one member's steps/s = members run one after another
    value = steps_b * members / wall_b
    return {
        "workload": f"{members} members x 4 tracers on the zoom tree (nt={members * 4}), FP64",
        "value": round(value, 2),
        "unit": "member-steps/s",
        "sequential_member_steps_per_s": round(sequential, 2),
        "speedup_vs_sequential": round(value / sequential, 3),
        "ms_per_interval": round(wall_b / intervals * 1e3, 3),
        "intervals": intervals,
    }


def adjoint_section(met):
    """The gradient of the 500-sample column-average misfit over 24 hourly intervals at
    glb3x2x34 with respect to the emission field and the initial tracers, through the
    torch.autograd.Function (checkpointed forward, replayed intervals, device adjoint)."""
    import torch

    from pydyn_amd import adjoint as AD
    from pydyn_amd import device as D

    window = AD.Window(met, 24)
    i, j, obs = AD.samples(met.g)
    st = met.state()
    init = [D.upload(st[k]) for k in ("rm", "rxm", "rym", "rzm")]
    emis = D.upload(met.emis)
    times = []
    for _ in range(2):  # the first pass warms up
        x = [t.clone().requires_grad_(True) for t in init]
        e = emis.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = AD.transport(window, *x, e)
        L = AD.loss(out[0], window.m_final, i, j, obs)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.backward()
        torch.cuda.synchronize()
        times.append((t1 - t0, time.perf_counter() - t1))
    fwd, bwd = times[-1]
    return {
        "workload": "24 x 3600 s of glb3x2x34, loss over 500 seeded columns, d/d(emission 120x90, initial tracers)",
        "forward_s": round(fwd, 4),
        "backward_s": round(bwd, 4),
        "gradients_per_hour": round(3600.0 / (fwd + bwd), 1),
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
    if a.only:
        out = {"zoom": zoom_section, "batched": lambda: batched_section(a.members),
               "adjoint": lambda: adjoint_section(met)}[a.only]()
        print(json.dumps({a.only: out}), flush=True)
        return 0
    reg = Region(met)
    for _ in range(a.warmup):
        reg.interval()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim = sum(ndyn * n for ndyn, n in reg.run(a.steps))  # one read of the device for the K intervals
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
        "roofline": roofline(reg, a.profile_intervals),
    }
    if not a.no_zoom:
        line["zoom"] = zoom_section()
    if not a.no_batched:
        line["batched"] = batched_section(a.members)
    if not a.no_adjoint:
        line["adjoint"] = adjoint_section(met)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(met, a.cpu_intervals)
    print(json.dumps(line), flush=True)
    if a.copy_calibration:
        src = torch.ones(a.copy_calibration << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
    return 0


if __name__ == "__main__":
    sys.exit(main())
