"""Tile shapes of the line, column and many-tracer kernels, measured on the GPU.

    python -m pydyn_amd.tune [--workload region|batched|all] [--intervals 6] [--write]

Each knob of tm5_ctx_set_tile (device.TILES) is tuned on the workload whose kernels it shapes:

  region   the plain step on one global region (glb3x2 x 34, one tracer): the y-line tile, the
           dynam0 and m2phlb1 tiles;
  batched  64 members x 4 tracers on the production zoom tree (nt = 256): the z-line tile (z runs
           on its own in a tree), the y tile of the many-tracer sweeps, the fewest tracers a
           tracer range holds, and the apply tile of the split column form.

For each knob in turn every candidate value is set on the context and the workload is built
afresh and run over the same queued intervals (two to warm up, then the timed ones, so that every
candidate meets the same Check_CFL cuts) with per-kernel timing (the library's events around each
launch); the value whose kernels take the least GPU time per dynamic step is kept before the next
knob is tried.
--write stores the result in pydyn_amd/tiles.json, which every new context applies.  A tile shape
changes how the work is split into workgroups, never an operation or its order, so the results
are the same bits for every value (the GPU tests run with the file's, and with other shapes).
"""

import argparse
import json
import sys

# workload -> knob -> (candidate values, the kernels it shapes; None: every kernel of the step)
CANDIDATES = {
    "region": {
        "y_lines": ((2, 3, 4, 5, 6, 7, 8, 10, 11, 16), ("k_y_steps<false>", "k_y_steps<true>", "k_y_mass1")),
        # a step's first x (or the launch with the step before's last) makes the step's fluxes
        "dynam0": ((4, 8, 16, 32, 64), ("k_dynam0", "k_x<false, true>", "k_xx")),
        "m2phlb": ((4, 8, 16, 32, 64), ("k_m2phlb1",)),
    },
    "batched": {
        "z_lines": ((4, 8, 16, 30, 32), None),
        "y_many": ((4, 6, 8, 11, 16), None),
        "range_min": ((4, 8, 16, 32, 64), None),
        "apply": ((8, 16, 32), None),
    },
}


def _region(members):
    from pydyn_amd import meteo
    from pydyn_amd.region import Region

    met = meteo.Meteo(meteo.grid("glb3x2"), seed=1)

    def make():
        reg = Region(met)
        return reg.run, reg.close

    return make


def _batched(members):
    from pydyn_amd import meteo
    from pydyn_amd.zoom.run import ZoomRun

    p, c = meteo.zoom("c2", nt=4)
    ens = meteo.Ensemble(p, members)

    def make():
        run = ZoomRun(p, c, ensemble=ens)
        return run.run, run.close

    return make


WORKLOADS = {"region": _region, "batched": _batched}


def _time(make, ctx, kernels, intervals):
    """GPU time per dynamic step of intervals 2 .. intervals + 1 of a fresh run (the kernels
    listed, or all)."""
    run, close = make()
    try:
        run(2)  # warm up (and capture the step graphs)
        ctx.profile_begin()
        out = run(intervals)
        prof = ctx.profile_end()
    finally:
        close()
    steps = sum(n for _, n in out)
    return sum(s for k, (_, s) in prof.items() if kernels is None or k in kernels) / steps


def tune(workload="region", intervals=6, members=64, log=print):
    from pydyn_amd import device as D

    make = WORKLOADS[workload](members)
    ctx = D.context()
    best = {}
    for knob, (values, kernels) in CANDIDATES[workload].items():
        times = {}
        for v in values:
            try:
                ctx.set_tile(knob, v)
            except Exception:  # outside the knob's range for this grid
                continue
            times[v] = min(_time(make, ctx, kernels, intervals) for _ in range(2))
            if times[v] == 0.0:  # the workload launches none of these kernels: keep the default
                times.pop(v)
        if not times:
            continue
        best[knob] = min(times, key=times.get)
        ctx.set_tile(knob, best[knob])
        log(f"{knob:9s} " + "  ".join(f"{v}: {t * 1e6:9.2f} us" for v, t in times.items()) + f"  -> {best[knob]}"
            "  (GPU time a step)")
    return best


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--workload", choices=("region", "batched", "all"), default="all")
    p.add_argument("--intervals", type=int, default=6)
    p.add_argument("--members", type=int, default=64, help="members of the batched workload (4 tracers each)")
    p.add_argument("--write", action="store_true", help="store the result in pydyn_amd/tiles.json")
    p.add_argument("--out", help="store the result in this file instead")
    a = p.parse_args(argv)
    import pathlib

    from pydyn_amd import device as D

    D.TILES_FILE_ENABLED = False  # measure from the library's defaults
    best, names = {}, {"region": "glb3x2 x 34, 1 tracer", "batched": f"c2 tree, {a.members} members x 4 tracers"}
    done = {}
    for w in ("region", "batched") if a.workload == "all" else (a.workload,):
        print(f"== {w}: {names[w]}", flush=True)
        best.update(tune(w, a.intervals, a.members, log=lambda s: print(s, flush=True)))
        done[w] = names[w]
    out = {"workloads": done, "device": "MI355X (gfx950)", "tiles": best}
    print(json.dumps(out))
    path = pathlib.Path(a.out) if a.out else D.TILES_FILE if a.write else None
    if path is not None:
        if a.workload != "all" and path.exists():  # keep the other workload's knobs
            old = json.loads(path.read_text())
            out["workloads"] = {**old.get("workloads", {}), **done}
            out["tiles"] = {**old.get("tiles", {}), **best}
        path.write_text(json.dumps(out, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
