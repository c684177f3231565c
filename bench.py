"""Throughput of the TM5 transport step on one MI355X: simulated years per hour at glb3x2x34.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step here is one meteo interval of 3600 s on the glb3x2 global grid (120 x 90 x 34, one CH4
tracer, FP64): Check_CFL (the mass-only rehearsal that may cut ndyn), then its n steps of ndyn,
each Setup_MassFlow + advection x/y/z with sub-cycling, convection with Kzz diffusion, CH4
injection and chemistry, m2phlb1 -- the state resident in HBM, inputs synthetic (seeded,
pydyn_amd.meteo).  Every interval is bit-identical to the reference (tests/test_gpu_parity.py).

The timed K intervals are the first K of the 30-day run (interval 0 onwards, on a region built
for the timing; the W warm-up intervals run on another region of the same meteo).  With the
default K = 720 that is the whole 30-day configuration, Check_CFL's cut intervals included; with
a smaller K the line also carries the whole 30 days as `full_config`.

value = simulated seconds advanced / wall seconds x 3600 / (365 x 86400), the reference's
years/hour (pydyn/gpu/bench.py:26, 262) with the ndyn Check_CFL actually chose.

Also reported, in the same line: the step's algorithmic HBM bytes over its time ("hbm"), the zoom
tree's rate ("zoom", 7 days), the reference's own benchmarked tree ("published_tree"), the
64-member x 4-tracer batched form on the zoom tree against one member's plain run ("batched"),
the cost of one gradient over 24 hourly intervals through the adjoint ("adjoint"), convec with the
LAPACK kind's inverse against the Crout kind ("lapack"), the dominant
kernel's roofline and the step's other kernels (their algorithmic bytes or FP64 operations per
launch over their average launch time, from HIP events the library records around each launch on
its stream, over intervals run after the timed ones) and the oracle (NumPy, one core) timed on a
bounded sample of the same workload on this host.
"""

import argparse
import contextlib
import gc
import json
import os
import sys
import time

SECONDS_A_YEAR = 365.0 * 86400.0
METRIC = "simulated years/hour at glb3x2x34 (FP64 bit-exact); 64-member member-steps/s; HBM GB/s"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix); it counts an FMA as two operations
C1_INTERVALS = 720  # the 30-day configuration


def parse():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=C1_INTERVALS, help="timed intervals of 3600 s (from t = 0)")
    p.add_argument("--warmup", type=int, default=4)
    p.add_argument("--grid", default="glb3x2")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-intervals", type=int, default=24)
    p.add_argument("--profile-intervals", type=int, default=12, help="intervals timed per kernel after the run")
    p.add_argument("--no-zoom", action="store_true", help="skip the zoom trees' rates")
    p.add_argument("--no-batched", action="store_true", help="skip the 64-member batched rate")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint window's cost")
    p.add_argument("--members", type=int, default=64)
    p.add_argument("--only", choices=("zoom", "published_tree", "batched", "adjoint", "lapack"),
                   help="run that section alone and print its object (for profiling one workload)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="end with three plain device copies of MIB MiB (a known size for rocprofv3 --pmc)")
    return p.parse_args()


# ------------------------------------------------------------------ algorithmic work
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


def _shapes(im, jm, lm):
    cells, cols = im * jm * lm, im * jm
    return cells, cols, (im + 1) * jm * lm, im * (jm + 1) * lm, im * jm * (lm + 1)


def operator_bytes(im, jm, lm, lmax, nt):
    """Algorithmic HBM bytes of each operator of a step on an im x jm x lm region with nt tracers:
    one read and one write of every array the operator changes, one read of every array it only
    reads (sub-cycles stay on chip)."""
    cells, cols, fx, fy, fz = _shapes(im, jm, lm)
    q = 4 * nt  # rm, rxm, rym, rzm of every tracer
    return {
        "x": 8 * (2 * cells * (1 + q) + 2 * fx),  # m and moments in and out, am scaled in place
        "y": 8 * (2 * cells * (1 + q) + 2 * fy),
        "z": 8 * (2 * cells * (1 + q) + fz),  # cm read (restored exactly)
        "convec": 8 * (2 * cells * q + cells + 5 * cols * lmax) + 4 * 2 * cols,  # + m, 4 fluxes, dkg, 2 levels
        "sources": 8 * (cols + 2 * 2 * cols),  # emission; rm, rzm of the emitting level
        "chemistry": 8 * (4 * cells + 2 * cells * q),  # the four loss rates; the moments
        "m2phlb1": 8 * (cells + cols * (lm + 2)),  # m in; sp, phlb out
        "dynam0": 8 * (2 * fx + 2 * fy + fz),  # pu, pv in; am, bm, cm out
    }


def step_bytes(im, jm, lm, lmax, nt):
    """A step's algorithmic bytes: x, y, z, convection, sources, chemistry and m2phlb1 twice
    ("xyzvsc m2phlb1 csvzyx m2phlb1"), Setup_MassFlow's dynam0 once."""
    b = operator_bytes(im, jm, lm, lmax, nt)
    return 2 * sum(v for k, v in b.items() if k != "dynam0") + b["dynam0"]


def tree_step_bytes(zoom, lmax, nt):
    """A parent step of a tree: every region's step bytes, a child's once per its time refinement
    (the exchange and coarsening, a few cells per wall, not counted)."""
    total = 0
    for r in zoom.regions():
        g = zoom.g(r)
        total += g.tref * step_bytes(g.im, g.jm, zoom.lm, lmax, nt)
    return total


def kernel_units(g, nt):
    """Algorithmic work per launch of the region step's kernels at grid g: (bound, amount, unit)."""
    cells, cols, fx, fy, fz = _shapes(g.im, g.jm, g.lm)
    q = 4 * nt
    return {
        "k_column<%d>" % g.lmax: ("fp64_valu", cols * column_flops(g.lmax, nt), "flop"),
        "k_x<false, false>": ("hbm", 8 * (2 * cells * (1 + q) + 2 * fx), "B"),
        # a step's first x making its fluxes: am from pu (read in place of am), and dynam0's tiles
        # reading pu, pv and writing bm, cm
        "k_x<false, true>": ("hbm", 8 * (2 * cells * (1 + q) + 2 * fx + (fx + fy) + (fy + fz)), "B"),
        # a step's last x and the next step's first: the row's m and moments read and written once,
        # am read, pu read, am written, and the dynam0 tiles as above
        "k_xx": ("hbm", 8 * (2 * cells * (1 + q) + 3 * fx + (fx + fy) + (fy + fz)), "B"),
        "k_y_steps<false>": ("hbm", 8 * (2 * cells * (1 + q) + fy), "B"),
        "k_z_steps<false>": ("hbm", 8 * (2 * cells * (1 + q) + fz), "B"),
        "k_x<true, false>": ("hbm", 8 * (2 * cells + 2 * fx), "B"),
        "k_y_steps<true>": ("hbm", 8 * (2 * cells + fy), "B"),
        "k_z_steps<true>": ("hbm", 8 * (2 * cells + fz), "B"),
        "k_z_search": ("hbm", 8 * (cells + fz), "B"),
        "k_y_search": ("hbm", 8 * (cells + fy), "B"),
        "k_dynam0": ("hbm", 8 * (2 * fx + 2 * fy + fz), "B"),
        "k_m2phlb1": ("hbm", 8 * (cells + cols * (g.lm + 2)), "B"),
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


def roofline(met, intervals, skip=4):
    """Per-kernel GPU time from events the library records around every launch on its stream, over
    `intervals` intervals of a region of its own from t = 0 after `skip` (C1's intervals 4..15: no
    Check_CFL cut there, so no plan misses and every launch does its work -- after a miss the
    queued launches skip in about a microsecond, which would pull the averages down): the
    roofline of the kernel with the largest total, and every kernel's achieved rate."""
    from pydyn_amd import device as D
    from pydyn_amd.region import Region

    if intervals < 1:
        return None, None
    reg = Region(met)
    reg.run(skip)
    ctx = D.context()
    ctx.profile_begin()
    out = reg.run(intervals)
    prof = ctx.profile_end()
    cut = sum(1 for _, n in out if n != 1)
    units = kernel_units(reg.g, reg.rm.shape[3])
    reg.close()
    total = sum(s for _, s in prof.values())
    kernels = []
    for name, (count, secs) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
        avg = secs / count
        row = {"kernel": name, "launches": count, "launch_us": round(avg * 1e6, 2), "share_of_gpu_time": round(secs / total, 3)}
        if name in units:
            bound, amount, unit = units[name]
            row["algorithmic_per_launch"], row["algorithmic_unit"] = amount, unit
            if bound == "hbm":
                row["achieved_GBs"] = round(amount / avg / 1e9, 1)
                row["frac_of_hbm_peak"] = round(amount / avg / 1e9 / HBM_PEAK_GBS, 4)
            else:
                row["achieved_TFLOPs"] = round(amount / avg / 1e12, 3)
        traffic = traffic_of(name)
        if traffic is not None:
            row["traffic_B"] = traffic
            row["traffic_GBs"] = round(traffic / avg / 1e9, 1)
        kernels.append(row)
    top = kernels[0]
    if top["kernel"] not in units:
        return {"kernel": top["kernel"], "bound": None, "launch_us": top["launch_us"]}, kernels
    bound, amount, unit = units[top["kernel"]]
    avg = top["launch_us"] * 1e-6
    if bound == "hbm":
        achieved, peak, runit = amount / avg / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        achieved, peak, runit = amount / avg / 1e12, FP64_PEAK_TFLOPS, "TFLOP/s"
    out = {
        "kernel": top["kernel"],
        "bound": bound,
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": runit,
        "frac": round(achieved / peak, 4),
        "traffic": traffic_of(top["kernel"]),
        "algorithmic_per_launch": amount,
        "algorithmic_unit": unit,
        "launch_us": top["launch_us"],
        "launches": top["launches"],
        "share_of_gpu_time": top["share_of_gpu_time"],
        "timed_intervals": f"{skip}..{skip + intervals - 1} of a region of its own ({cut} cut)",
    }
    if bound == "fp64_valu":
        out["frac_no_fma"] = round(achieved / (peak / 2.0), 4)
        out["note"] = ("FP64 vector adds, multiplies and divisions, no FMA (bit-exactness) and no matrix core; "
                       "peak is the published FP64 figure, which counts an FMA as two operations: frac_no_fma "
                       "is against half of it")
    return out, kernels


# ------------------------------------------------------------------ sections
@contextlib.contextmanager
def _no_gc():
    """A timed region without Python's cyclic garbage collector: it is collected before and held
    off until after (as timeit does).  A full collection with torch loaded takes tens of ms, and
    one landing inside a 0.13 s window would be timed as GPU work."""
    gc.collect()
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def _years(sim, wall):
    return sim / wall * 3600.0 / SECONDS_A_YEAR


def _queued(run, count):
    """count intervals of a ZoomRun queued on the native driver: (wall s, [(ndyn, n)])."""
    import torch

    with _no_gc():
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = run.run(count)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
    return wall, out


def _launches(run, intervals):
    """Kernel launches per interval over `intervals` more queued intervals (the library's events;
    graphs are not replayed while they are recorded)."""
    from pydyn_amd import device as D

    ctx = D.context()
    ctx.profile_begin()
    run.run(intervals)
    prof = ctx.profile_end()
    return round(sum(c for c, _ in prof.values()) / intervals, 1)


def tree_section(name, workload, days=7, warmup=4, published=None):
    """A zoom tree, one CH4 tracer, queued on the native driver: after `warmup` intervals (which
    capture the step graphs), 24 intervals that Check_CFL does not cut, then `days` days."""
    from pydyn_amd import meteo
    from pydyn_amd.zoom.run import ZoomRun

    p, c = meteo.zoom(name)
    run = ZoomRun(p, c, redgrid=meteo.redgrid(name))
    run.run(warmup)
    wall_u, out_u = _queued(run, 24)
    wall, out = _queued(run, 24 * days)
    sim = sum(a * b for a, b in out)
    uncut = sum(1 for _, n in out_u if n == 1) == len(out_u)
    bytes_step = tree_step_bytes(run.zoom, p.g.lmax, 1)
    steps = sum(n for _, n in out)
    rec = {
        "workload": workload,
        "value": round(_years(sim, wall), 3),
        "unit": "simulated years/hour",
        "intervals": len(out),
        "cut_intervals": sum(1 for _, n in out if n != 1),
        "ms_per_interval": round(wall / len(out) * 1e3, 3),
        "uncut_window": {
            "intervals": len(out_u),
            "all_uncut": uncut,
            "ms_per_interval": round(wall_u / len(out_u) * 1e3, 3),
            "value": round(_years(sum(a * b for a, b in out_u), wall_u), 3),
            "launches_per_interval": _launches(run, 2) if uncut else None,
        },
        "hbm": {
            "algorithmic_bytes_per_parent_step": bytes_step,
            "achieved_GBs": round(bytes_step * steps / wall / 1e9, 1),
            "frac_of_peak": round(bytes_step * steps / wall / 1e9 / HBM_PEAK_GBS, 4),
        },
        "tally": run.native().tally(),
    }
    if published:
        rec["published"] = published
        rec["vs_published"] = {k: round(v["ms_per_interval"] / rec["uncut_window"]["ms_per_interval"], 2)
                               for k, v in published.items()}  # fmt: skip
    run.close()
    return rec


def zoom_section():
    return tree_section("c2", "glb3x2x34 + child 120x70x34 (xref 3, yref 2, tref 2), two-way exchange, 1 CH4 tracer, "
                        "7 days, queued on the native tree driver (checked steps replayed from HIP graphs)")


def published_tree_section():
    """The tree the reference publishes its device timings on (pydyn/gpu/bench.py:28-41):
    glb3x2 on its reduced grid + nam1x1 72 x 44 (xref 3, yref 2, tref 2), 34 levels, convection
    over 26, one tracer -- on this build's synthetic meteo (the reference's interval: Check_CFL +
    Setup_MassFlow + step, as the uncut window here)."""
    published = {  # ms per interval, replayed as a graph (docs/gpu_port.md:196-202, 512-518)
        "A100_80GB": {"ms_per_interval": 6.8, "years_per_hour": 60.1},
        "MI100": {"ms_per_interval": 13.9, "years_per_hour": 29.6},
    }
    return tree_section("pub", "glb3x2 (reduced grid) x 34 + nam1x1 72x44 (xref 3, yref 2, tref 2), 1 CH4 tracer, "
                        "7 days, queued on the native tree driver", published=published)


def batched_section(members, intervals=24):
    """members x 4 tracers on the zoom tree in one run (each member's emission scaled, each
    bit-identical to its own plain run) against one member's plain run: member-steps/s, a step
    being one parent step of the tree; both queued on the native driver."""
    import torch

    from pydyn_amd import meteo
    from pydyn_amd.zoom.run import ZoomRun

    p, c = meteo.zoom("c2", nt=4)
    ens = meteo.Ensemble(p, members)
    plain = ZoomRun(p, c, ensemble=ens.subset([0]))
    plain.run(4)
    wall_p, out_p = _queued(plain, 2 * intervals)
    plain.close()
    del plain
    torch.cuda.empty_cache()
    batched = ZoomRun(p, c, ensemble=ens)
    batched.run(2)
    wall_b, out_b = _queued(batched, intervals)
    steps_p, steps_b = sum(n for _, n in out_p), sum(n for _, n in out_b)
    sequential = steps_p / wall_p  # one member's steps/s = members run one after another
    value = steps_b * members / wall_b
    bytes_step = tree_step_bytes(batched.zoom, p.g.lmax, ens.nt)
    rec = {
        "workload": f"{members} members x 4 tracers on the zoom tree (nt={members * 4}), FP64, queued on the native "
                    "tree driver (plans rehearsed and verified on the device, steps replayed from HIP graphs)",
        "value": round(value, 2),
        "unit": "member-steps/s",
        "sequential_member_steps_per_s": round(sequential, 2),
        "speedup_vs_sequential": round(value / sequential, 3),
        "ms_per_interval": round(wall_b / intervals * 1e3, 3),
        "intervals": intervals,
        "cut_intervals": sum(1 for _, n in out_b if n != 1),
        "hbm": {
            "algorithmic_bytes_per_parent_step": bytes_step,
            "achieved_GBs": round(bytes_step * steps_b / wall_b / 1e9, 1),
            "frac_of_peak": round(bytes_step * steps_b / wall_b / 1e9 / HBM_PEAK_GBS, 4),
        },
        "tally": batched.native().tally(),
    }
    batched.close()
    return rec


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
        with _no_gc():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out = AD.transport(window, *x, e)
            L = AD.loss(out[0], window.m_final, i, j, obs)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            L.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
        times.append((t1 - t0, t2 - t1))
    fwd, bwd = times[-1]
    return {
        "workload": "24 x 3600 s of glb3x2x34, loss over 500 seeded columns, d/d(emission 120x90, initial tracers)",
        "forward_s": round(fwd, 4),
        "backward_s": round(bwd, 4),
        "gradients_per_hour": round(3600.0 / (fwd + bwd), 1),
    }


def lapack_section(met, repeats=20):
    """convec over every column of glb3x2x34 (one tracer, dt 1800 s) with the LAPACK kind's
    inverse (DGETRF2 + DGETRI per column's 26 x 26 matrix, pydyn/lapack.py:43-66) against the
    Crout kind: each kind's kernels per launch, from the library's events."""
    import numpy as np
    import torch

    from pydyn_amd import device as D
    from pydyn_amd.gpu import kernels as K

    st = met.state()
    m = D.upload(st["m"])
    conv = {k: D.upload(np.ascontiguousarray(getattr(met, k))) for k in ("entu", "detu", "entd", "detd")}
    conv["dkg"] = K.dkg(D.upload(met.kvh), m, D.upload(met.gph), met.g.lmax)
    for k in ("cloud_top", "cloud_lfs"):
        conv[k] = D.upload(getattr(met, k).astype(np.int32)).to(torch.int32)
    q = [D.upload(st[k]) for k in ("rm", "rxm", "rym", "rzm")]
    ctx = D.context()
    out = {"workload": f"convec over {met.g.im} x {met.g.jm} columns of {met.g.lm} levels (lmax {met.g.lmax}), "
                       "1 tracer, FP64, per launch"}  # fmt: skip
    for kind in ("lapack", "crout"):
        for _ in range(3):
            K.convec(m, *q, conv, 1800.0, met.g.lmax, inverse=kind)
        torch.cuda.synchronize()
        ctx.profile_begin()
        for _ in range(repeats):
            K.convec(m, *q, conv, 1800.0, met.g.lmax, inverse=kind)
        prof = ctx.profile_end()
        out[kind] = {k: round(v[1] / v[0] * 1e6, 2) for k, v in sorted(prof.items())}
        out[kind]["total_us"] = round(sum(v[1] for v in prof.values()) / repeats * 1e6, 2)
    out["unit"] = "us per launch"
    return out


def cpu_baseline(met, intervals):
    from oracle.driver import Run

    run = Run(met)
    sim = 0
    with _no_gc():  # as the device's timed regions
        t = time.perf_counter()
        for _ in range(intervals):
            ndyn, n = run.interval()
            sim += ndyn * n
        wall = time.perf_counter() - t
    return {
        "value": round(_years(sim, wall), 6),
        "unit": "simulated years/hour",
        "cores": 1,
        "kind": "port",
        "sample": f"{intervals} intervals ({sim} simulated s) of glb3x2x34 from t=0, the NumPy oracle, "
        f"single-threaded (os.cpu_count()={os.cpu_count()}, affinity {len(os.sched_getaffinity(0))})",
        "wall_s": round(wall, 2),
    }


def _region_run(met, count):
    """A fresh region from t = 0, count intervals queued (one read of the device per batch):
    (wall s, [(ndyn, n)], the region)."""
    import torch

    from pydyn_amd.region import Region

    reg = Region(met)
    with _no_gc():
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = reg.run(count)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
    reg.check()
    return wall, out, reg


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
        out = {"zoom": zoom_section, "published_tree": published_tree_section,
               "batched": lambda: batched_section(a.members), "adjoint": lambda: adjoint_section(met),
               "lapack": lambda: lapack_section(met)}[a.only]()  # fmt: skip
        print(json.dumps({a.only: out}), flush=True)
        return 0
    warm = Region(met)
    warm.run(max(a.warmup, 1))  # code loaded and every buffer sized, on a region of its own
    warm.close()
    torch.cuda.synchronize()
    wall, out, reg = _region_run(met, a.steps)
    sim = sum(ndyn * n for ndyn, n in out)
    steps = sum(n for _, n in out)
    g = met.g
    b_step = step_bytes(g.im, g.jm, g.lm, g.lmax, met.nt)
    line = {
        "metric": METRIC,
        "value": round(_years(sim, wall), 3),
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
            "workload": f"{a.grid} x {g.lm} plain FP64 interval (Check_CFL + steps), 1 CH4 tracer, intervals "
                        f"0..{a.steps - 1} of the 30-day run",
            "grid": [g.im, g.jm, g.lm],
            "lmax": g.lmax,
            "nt": met.nt,
            "interval_s": 3600,
            "simulated_s": sim,
            "dynamic_steps": steps,
            "cut_intervals": sum(1 for _, n in out if n != 1),
        },
        "ms_per_dynamic_step": round(wall / steps * 1e3, 4),
        "hbm": {
            "algorithmic_bytes_per_step": b_step,
            "achieved_GBs": round(b_step * steps / wall / 1e9, 1),
            "frac_of_peak": round(b_step * steps / wall / 1e9 / HBM_PEAK_GBS, 4),
        },
    }
    if a.steps < C1_INTERVALS:
        wall_f, out_f, reg_f = _region_run(met, C1_INTERVALS)
        reg_f.close()
        sim_f, steps_f = sum(x * y for x, y in out_f), sum(n for _, n in out_f)
        line["full_config"] = {
            "intervals": C1_INTERVALS,
            "cut_intervals": sum(1 for _, n in out_f if n != 1),
            "value": round(_years(sim_f, wall_f), 3),
            "ms_per_interval": round(wall_f / C1_INTERVALS * 1e3, 4),
            "hbm_achieved_GBs": round(b_step * steps_f / wall_f / 1e9, 1),
        }
    reg.close()
    line["roofline"], line["kernels"] = roofline(met, a.profile_intervals)
    if not a.no_zoom:
        line["zoom"] = zoom_section()
        line["published_tree"] = published_tree_section()
    if not a.no_batched:
        line["batched"] = batched_section(a.members)
    if not a.no_adjoint:
        line["adjoint"] = adjoint_section(met)
        line["lapack"] = lapack_section(met)
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
