"""Tile shapes of the line and column kernels, measured on the GPU.

    python -m pydyn_amd.tune [--grid glb3x2] [--intervals 6] [--write]

For each knob of tm5_ctx_set_tile (device.TILES), in turn, every candidate value is set on the
context and the region's queued intervals are run with per-kernel timing (the library's events
around each launch); the value whose kernels take the least GPU time is kept before the next
knob is tried.  --write stores the result in pydyn_amd/tiles.json, which every new context
applies.  A tile shape changes how the work is split into workgroups, never an operation or its
order, so the results are the same bits for every value (the GPU tests run with the file's).
"""

import argparse
import json
import sys

# knob -> (candidate values, the kernels it shapes)
CANDIDATES = {
    "y_lines": ((2, 4, 6, 8, 11, 16), ("k_y_steps<false>", "k_y_steps<true>", "k_y_mass1")),
    "z_lines": ((4, 8, 16, 30, 32), ("k_z_steps<false>", "k_z_steps<true>", "k_z_mass1")),
    "dynam0": ((4, 8, 16, 32, 64), ("k_dynam0",)),
    "m2phlb": ((4, 8, 16, 32, 64), ("k_m2phlb1",)),
}


def _time(reg, ctx, kernels, intervals):
    ctx.profile_begin()
    reg.run(intervals)
    prof = ctx.profile_end()
    return sum(prof[k][1] for k in kernels if k in prof)


def tune(grid="glb3x2", intervals=6, log=print):
    from pydyn_amd import device as D
    from pydyn_amd import meteo
    from pydyn_amd.region import Region

    met = meteo.Meteo(meteo.grid(grid), seed=1)
    reg = Region(met)
    ctx = D.context()
    reg.run(2)  # warm up
    best = {}
    for knob, (values, kernels) in CANDIDATES.items():
        times = {}
        for v in values:
            try:
                ctx.set_tile(knob, v)
            except Exception:  # outside the knob's range for this grid
                continue
            reg.run(1)
            times[v] = min(_time(reg, ctx, kernels, intervals) for _ in range(2))
            if times[v] == 0.0:  # the grid's run launches none of these kernels: keep the default
                times.pop(v)
        if not times:
            continue
        best[knob] = min(times, key=times.get)
        ctx.set_tile(knob, best[knob])
        log(f"{knob:8s} " + "  ".join(f"{v}: {t / intervals * 1e6:8.2f} us" for v, t in times.items()) + f"  -> {best[knob]}")
    return best


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--grid", default="glb3x2")
    p.add_argument("--intervals", type=int, default=6)
    p.add_argument("--write", action="store_true", help="store the result in pydyn_amd/tiles.json")
    p.add_argument("--out", help="store the result in this file instead")
    a = p.parse_args(argv)
    import pathlib

    from pydyn_amd import device as D

    D.TILES_FILE_ENABLED = False  # measure from the library's defaults
    best = tune(a.grid, a.intervals, log=lambda s: print(s, flush=True))
    out = {"grid": a.grid, "device": "MI355X (gfx950)", "tiles": best}
    print(json.dumps(out))
    path = pathlib.Path(a.out) if a.out else D.TILES_FILE if a.write else None
    if path is not None:
        path.write_text(json.dumps(out, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
