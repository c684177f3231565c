"""Write the golden vectors of the TM5 transport step from the reference implementation itself.

Run in the build container only (the reference is not on the GPU machines):

    python tests/golden/make_golden.py ops            # operator cases at tiny shapes (npz)
    python tests/golden/make_golden.py run tiny 48    # whole single-region runs: sha1 per interval
    python tests/golden/make_golden.py run glb6x4 24
    python tests/golden/make_golden.py run glb3x2 720 # long: written interval by interval
    python tests/golden/make_golden.py zoom tiny2 48  # a global region with one child (zoom tree)
    python tests/golden/make_golden.py zoom c2e 40    # 2 members x 4 tracers on the production tree
    python tests/golden/make_golden.py chem           # CH4 loss rates and decay (npz)
    python tests/golden/make_golden.py kzz            # Calc_Kzz: dkg and its intermediates (npz)

It imports the reference package from /root/reference (read-only) and routes the reference's
libm through the oracle's fdlibm library with the reference's own `ops.use_libm`, so that its
CH4 decay factors use the same exp as the device.  Inputs come from pydyn_amd.meteo (seeded) or
from the reference's own synthetic generator (pydyn.synth) at tiny shapes; the fixtures hold
those inputs and the reference's outputs, nothing else.

The reference's emission module cannot be imported here (it imports netCDF4 at module top), so
the injected mass uses the three-operation formula of pydyn/emis.py:52-54 restated below
(`_inject`); month2dt and XMCH4 are pinned separately by the reference tests' known answers.
"""

import hashlib
import json
import os
import pathlib
import sys

import numpy as np

REPO = pathlib.Path(__file__).resolve().parents[2]
GOLDEN = REPO / "tests" / "golden"
REFERENCE = pathlib.Path("/root/reference")

# the reference's libm route compiles a C loop into its cache: the oracle's (git-ignored) build folder
os.environ.setdefault("XDG_CACHE_HOME", str(REPO / "oracle" / "_build" / "reference_cache"))
sys.path.insert(0, str(REPO))
sys.path.insert(1, str(REFERENCE))

from pydyn import cfl, chem, diffusion, kernels, ops, synth  # noqa: E402
from pydyn import zoom as Z  # noqa: E402
from pydyn.redgrid import RedGrid  # noqa: E402

from oracle import libm as oracle_libm  # noqa: E402
from pydyn_amd import meteo  # noqa: E402

XMCH4 = 12.01115 + 1.0079 * 4.0
SEC_MONTH = 30.0 * 86400.0


def _use_twin():
    oracle_libm.build()
    ops.use_libm(str(oracle_libm.LIB_PATH))


def _inject(emis, ndyn, tref=1):
    month2dt = (float(ndyn) / (2 * tref)) / SEC_MONTH  # pydyn/emis.py:65-68
    return ((emis * month2dt) * (XMCH4 / XMCH4)) * 1.0  # pydyn/emis.py:52-54


def sha(a):
    return hashlib.sha1(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()[:20]


# ------------------------------------------------------------------ operator cases
def _state(rng, shape):
    s = synth.state(rng, *shape)
    return {k: getattr(s, k).copy() for k in ("m", "rm", "rxm", "rym", "rzm")}


def ops_cases():
    out = {}
    shape = (8, 6, 5, 2)

    def put(case, **arrays):
        for k, v in arrays.items():
            out[f"{case}.{k}"] = np.asarray(v)

    # advection, with and without sub-cycling and limiter
    for axis, (kern, idx) in {"x": (kernels.advect_x, 0), "y": (kernels.advect_y, 1), "z": (kernels.advect_z, 2)}.items():
        for cfl_, limits, seed in ((0.4, True, 1), (2.5, True, 2), (0.9, False, 3), (3.5, True, 4)):
            rng = np.random.default_rng(100 * seed + idx)
            st = _state(rng, shape)
            if seed == 4:  # slopes as large as the limiter allows, some beyond it
                for k in ("rxm", "rym", "rzm"):
                    st[k] = 1.3 * st["rm"] * rng.choice([-1.0, 1.0], st["rm"].shape)
            flux = synth.fluxes(rng, st["m"], cfl=cfl_)[idx].copy()
            case = f"adv_{axis}_{seed}"
            put(case, **{f"in_{k}": v.copy() for k, v in st.items()}, in_flux=flux.copy(), limits=limits)
            stats = {}
            kern(st["m"], st["rm"], st["rxm"], st["rym"], st["rzm"], flux, limits, stats=stats)
            put(case, **{f"out_{k}": v for k, v in st.items()}, out_flux=flux, nloop_max=stats["nloop_max"])

    # dynam0
    rng = np.random.default_rng(7)
    im, jm, lm = 8, 6, 5
    pu = 1e5 * rng.uniform(-1, 1, (im + 1, jm, lm))
    pu[im] = pu[0]
    pv = 1e5 * rng.uniform(-1, 1, (im, jm + 1, lm))
    pv[:, 0] = pv[:, jm] = 0.0
    bt = np.linspace(1.0, 0.0, lm + 1) ** 1.5
    am, bm, cm = kernels.fluxes_from_meteo(pu, pv, bt, 900.0)
    put("dynam0", pu=pu, pv=pv, bt=bt, dt=900.0, am=am, bm=bm, cm=cm)

    # convection: the reference generator's variants, with and without a cloud above lmax
    for seed, (name, kw, lmax) in enumerate((
        ("conv_a", {}, 5),
        ("conv_b", {"strong_detu": True}, 5),
        ("conv_c", {"lfs_above_top": True}, 4),
        ("conv_d", {"top_max": 2}, 5),
    )):
        rng = np.random.default_rng(500 + seed)
        st = _state(rng, shape)
        c = synth.convection(rng, st["m"], lmax, 900.0, **kw)
        put(name, **{f"in_{k}": v.copy() for k, v in st.items()}, **{f"c_{k}": v for k, v in c.items()}, lmax=lmax)
        kernels.convec(st["m"], st["rm"], st["rxm"], st["rym"], st["rzm"], dt=900.0, lmax=lmax, **c)
        put(name, **{f"out_{k}": v for k, v in st.items()})
        ncol = shape[0] * shape[1]
        conv1 = kernels.convection_matrix(
            st["m"].reshape(ncol, -1)[:, :lmax], *(c[k].reshape(ncol, -1) for k in ("entu", "detu", "entd", "detd")),
            c["cloud_top"].ravel(), c["cloud_lfs"].ravel(), c["dkg"].reshape(ncol, -1), 900.0, lmax,
        )  # fmt: skip
        put(name, conv1=conv1, inverse=kernels.fastminv(conv1))

    # dkg
    rng = np.random.default_rng(8)
    kvh = rng.uniform(-5.0, 100.0, (im, jm, lm - 1))
    mm = 1e9 * rng.uniform(0.8, 1.2, (im, jm, lm))
    gph = np.cumsum(rng.uniform(100.0, 1000.0, (im, jm, lm + 1)), axis=-1)
    put("dkg", kvh=kvh, m=mm, gph=gph, lmax=4, dkg=diffusion.dkg(kvh, mm, gph, 4))

    # decay through the twin exp, and its application
    rng = np.random.default_rng(9)
    loss = rng.uniform(0.0, 1e-5, (im, jm, lm, 4))
    loss[..., 3] = 0.0
    loss[0, 0, 0] = [0.0, 0.0, 0.0, -1e-9]
    loss[1, 0, 0] = [2e-3, 1e-3, 5e-4, 0.0]  # arguments beyond ln2/2 take the reduction path
    put("decay", loss=loss, dt=1800.0, decay=chem.decay(loss, 1800.0))

    # mass-only kernels of Check_CFL
    for seed, c_ in ((11, 0.5), (12, 2.5), (13, 12.0)):
        rng = np.random.default_rng(seed)
        st = _state(rng, shape)
        am, bm, cm = synth.fluxes(rng, st["m"], cfl=c_)
        case = f"cflm_{seed}"
        put(case, m=st["m"], am=am, bm=bm, cm=cm)
        for axis, kern, flux, cap in (("x", cfl.dynamum, am, 10), ("y", cfl.dynamvm, bm, 1), ("z", cfl.dynamwm, cm, 1)):
            m = st["m"].copy()
            f = flux.copy()
            res = kern(m, f, max_nloop=cap) if axis != "z" else kern(m, f, iter_limit=cap)
            put(case, **{f"{axis}_ok": res is not None, f"{axis}_m": m, f"{axis}_flux": f})
    return out


# ------------------------------------------------------------------ whole single-region runs
def _interior(f, g):
    im, jm, lm = g.im, g.jm, g.lm
    return {
        "m": f["m"][1:im, 1:jm, 1:lm],
        **{k: f[k][1:im, 1:jm, 1:lm] for k in ("rm", "rxm", "rym", "rzm")},
        "am": f["am"][0:im, 1:jm, 1:lm],
        "bm": f["bm"][1:im, 1 : jm + 1, 1:lm],
        "cm": f["cm"][1:im, 1:jm, 0:lm],
        "sp": f["sp"][1:im, 1:jm, 1],
        "phlb": f["phlb"][1:im, 1:jm, 1 : lm + 1],
    }


def reference_run(met, n_intervals, every=None):
    """The single-region run of pydyn_amd.meteo's inputs through the reference's zoom driver:
    Check_CFL on the state, then its steps with Setup_MassFlow, sources and chemistry
    (pydyn/gpu/bench.py:154-160 and pydyn/forward.py:388-451).  every(it, t0, ndyn, n, arrays)."""
    g, nt = met.g, met.nt
    im, jm, lm = g.im, g.jm, g.lm
    region = Z.Region(g.name, im, jm, xcyc=1, touch_sp=1, touch_np=1)
    zm = Z.Zoom([region], lm, g.at, g.bt)
    st = Z.ZoomState.zeros(zm, nt)
    f = st[1]
    init = met.state()
    f["m"][1:im, 1:jm, 1:lm] = init["m"]
    for k in ("rm", "rxm", "rym", "rzm"):
        f[k][1:im, 1:jm, 1:lm] = init[k]
    f["dxyp"][1:jm] = g.dxyp

    def put_fluxes(fields, pu, pv):
        fields["pu"][0:im, 1:jm, 1:lm] = pu
        fields["pv"][1:im, 1 : jm + 1, 1:lm] = pv

    put_fluxes(f, *met.fluxes(0.0))
    Z.setup_massflow(zm, st, 3600)
    Z.m2phlb1(zm, st, 1)
    dkg = diffusion.dkg(met.kvh, met.m, met.gph, g.lmax)
    conv = {
        1: {
            "entu": met.entu, "detu": met.detu, "entd": met.entd, "detd": met.detd, "dkg": dkg,
            "cloud_top": met.cloud_top, "cloud_lfs": met.cloud_lfs, "lmax": g.lmax,
        }
    }  # fmt: skip
    t = 0.0
    for it in range(n_intervals):
        t0 = t

        def interpolate(zoom, s, i, ndyn, t0=t0):
            put_fluxes(s[1], *met.fluxes(t0 + (i - 0.5) * ndyn))

        ndyn, n, _ = Z.check_cfl_zoom(zm, st, 3600, 1, 3600, interpolate)
        for s in range(n):
            put_fluxes(f, *met.fluxes(t0 + s * ndyn + ndyn / 2.0))
            Z.setup_massflow(zm, st, ndyn)
            x = _inject(met.emis, ndyn)
            decay = chem.decay(met.loss, ndyn / 2.0)

            def sources(zoom, s_, region, x=x):
                Z.sources_region(zoom, s_, region, [(x, 1, 1)])

            def chemistry(zoom, s_, region, decay=decay):
                Z.chemistry_region(zoom, s_, region, decay, range(1, nt + 1))

            Z.step(zm, st, conv, ndyn, sources=sources, chemistry=chemistry)
        t = t0 + n * ndyn
        if every is not None:
            every(it, t0, ndyn, n, _interior(f, g))
    return _interior(f, g)


SEEDS = {"tiny": 3, "small": 5, "glb6x4": 1, "glb3x2": 1}


def run_golden(name, n_intervals):
    met = meteo.Meteo(meteo.grid(name), seed=SEEDS[name])
    path = GOLDEN / f"run_{name}_sha1.json"
    record = {"grid": name, "seed": SEEDS[name], "nt": met.nt, "intervals": []}

    def every(it, t0, ndyn, n, arrays):
        record["intervals"].append({"t0": t0, "ndyn": ndyn, "n": n, **{k: sha(v) for k, v in arrays.items()}})
        if it % 12 == 11 or it == n_intervals - 1:
            path.write_text(json.dumps(record, indent=0))
            print(f"{name}: interval {it + 1}/{n_intervals} t={t0 + n * ndyn:.0f} ndyn={ndyn}", flush=True)

    final = reference_run(met, n_intervals, every)
    if name in ("tiny", "small"):
        np.savez_compressed(GOLDEN / f"run_{name}_final.npz", **final)


# ------------------------------------------------------------------ zoom trees
ZFIELDS = ("m", "rm", "rxm", "rym", "rzm", "am", "bm", "cm", "pu", "pv", "sp", "phlb", "dxyp")


def reference_zoom_run(pmet, cmet, n_intervals, every=None, redgrid=None, ensemble=None):
    """A global region with one child on pydyn_amd.meteo's nested inputs through the reference's
    zoom path: the parent's footprint set to the child's sums (its update_parent), Check_CFL of
    the tree, then the steps with Setup_MassFlow; every region's `s` adds its emission (month2dt
    of its half step ndyn / (2 tref)) at level 1 of tracer 1 and its `c` applies the decay of its
    scope's loss rates over ndyn / (2 tref).  every(it, t0, ndyn, n, st)."""
    pg, c = pmet.g, cmet.c
    nt = pmet.nt if ensemble is None else ensemble.nt
    lm = pg.lm
    grid = None if redgrid is None else RedGrid(pg.im, pg.jm, *redgrid)
    glb = Z.Region("globe", pg.im, pg.jm, xcyc=1, touch_sp=1, touch_np=1, redgrid=grid)
    kid = Z.Region(c.name, c.im, c.jm, xbeg=c.xbeg, xend=c.xend, ybeg=c.ybeg, yend=c.yend, xref=c.xref,
                   yref=c.yref, tref=c.tref, parent=1)  # fmt: skip
    zm = Z.Zoom([glb, kid], lm, pg.at, pg.bt)
    assert (zm.ibeg[2], zm.jbeg[2], zm.iend[2], zm.jend[2]) == (c.ibeg, c.jbeg, c.iend, c.jend)
    st = Z.ZoomState.zeros(zm, nt)
    mets = {1: pmet, 2: cmet}
    conv, emis, loss = {}, {}, {}
    for r, met in mets.items():
        g, f = zm.g(r), st[r]
        init = met.state() if ensemble is None else ensemble.state(met)
        f["m"][1 : g.im, 1 : g.jm, 1:lm] = init["m"]
        for k in ("rm", "rxm", "rym", "rzm"):
            f[k][1 : g.im, 1 : g.jm, 1:lm] = init[k]
        f["dxyp"][1 : g.jm] = met.dxyp if r > 1 else pg.dxyp
        conv[r] = {
            "entu": met.entu, "detu": met.detu, "entd": met.entd, "detd": met.detd,
            "dkg": diffusion.dkg(met.kvh, met.m, met.gph, pg.lmax),
            "cloud_top": met.cloud_top, "cloud_lfs": met.cloud_lfs, "lmax": pg.lmax,
        }  # fmt: skip
        emis[r] = met.emis
        si, sj = Z.scope(zm, r)
        loss[r] = met.loss[si, sj]
    Z.update_parent(zm, st, 2)

    def put_fluxes(s_, t_sec):
        for r, met in mets.items():
            g = zm.g(r)
            pu, pv = met.fluxes(t_sec)
            s_[r]["pu"][0 : g.im, 1 : g.jm, 1:lm] = pu
            s_[r]["pv"][1 : g.im, 1 : g.jm + 1, 1:lm] = pv

    put_fluxes(st, 0.0)
    Z.setup_massflow(zm, st, 3600)
    Z.m2phlb(zm, st, 2)
    Z.m2phlb1(zm, st, 1)
    t = 0.0
    for it in range(n_intervals):
        t0 = t

        def interpolate(zoom, s_, i, ndyn, t0=t0):
            put_fluxes(s_, t0 + (i - 0.5) * ndyn)

        ndyn, n, _ = Z.check_cfl_zoom(zm, st, 3600, 1, 3600, interpolate)
        for s in range(n):
            put_fluxes(st, t0 + s * ndyn + ndyn / 2.0)
            Z.setup_massflow(zm, st, ndyn)

            def sources(zoom, s_, region, ndyn=ndyn):
                tref = zoom.g(region).tref
                if ensemble is None:
                    injections = [(_inject(emis[region], ndyn, tref), 1, 1)]
                else:  # every member's first tracer, its emission field scaled by its factor
                    injections = [(_inject(emis[region] * s, ndyn, tref), 1, t + 1)
                                  for t, s in enumerate(ensemble.escale()) if s != 0.0]  # fmt: skip
                Z.sources_region(zoom, s_, region, injections)

            def chemistry(zoom, s_, region, ndyn=ndyn):
                decay = chem.decay(loss[region], ndyn / (2.0 * zoom.g(region).tref))
                Z.chemistry_region(zoom, s_, region, decay, range(1, nt + 1))

            Z.step(zm, st, conv, ndyn, sources=sources, chemistry=chemistry)
        t = t0 + n * ndyn
        if every is not None:
            every(it, t0, ndyn, n, st)
    return zm, st


# ensembles: (zoom configuration, tracers per member, members)
ENSEMBLES = {"tiny2e": ("tiny2", 4, 4), "small2e": ("small2", 2, 3), "c2e": ("c2", 4, 2)}


def zoom_golden(name, n_intervals):
    ensemble = None
    if name in ENSEMBLES:
        base, ntr, members = ENSEMBLES[name]
        pmet, cmet = meteo.zoom(base, nt=ntr)
        ensemble = meteo.Ensemble(pmet, members)
    else:
        pmet, cmet = meteo.zoom(name)
    path = GOLDEN / f"zoom_{name}_sha1.json"
    record = {"config": name, "intervals": []}

    def every(it, t0, ndyn, n, st):
        arrays = {f"r{r}.{k}": sha(st[r][k].data) for r in (1, 2) for k in ZFIELDS}
        record["intervals"].append({"t0": t0, "ndyn": ndyn, "n": n, **arrays})
        if it % 6 == 5 or it == n_intervals - 1:
            path.write_text(json.dumps(record, indent=0))
            print(f"{name}: interval {it + 1}/{n_intervals} t={t0 + n * ndyn:.0f} ndyn={ndyn}", flush=True)

    redgrid = meteo.redgrid(ENSEMBLES[name][0] if name in ENSEMBLES else name)
    _, st = reference_zoom_run(pmet, cmet, n_intervals, every, redgrid, ensemble)
    if name in ("tiny2", "small2", "tiny2r", "tiny2e"):
        np.savez_compressed(GOLDEN / f"zoom_{name}_final.npz",
                            **{f"r{r}.{k}": st[r][k].data for r in (1, 2) for k in ZFIELDS})  # fmt: skip


def chem_cases():
    """pydyn.chem.loss_rates and decay on seeded columns: three cases -- the zonal columns alone,
    with the tropospheric Cl and the soil sink, and with the model's floor above the zonal
    columns' first edge (no fill from below).  Inputs and outputs, prefixed by the case."""
    out = {}
    im, jm, lm, nstrat, ntrop = 6, 5, 12, 8, 5
    for case, seed in (("plain", 1), ("trop", 2), ("high", 3)):
        rng = np.random.default_rng(40 + seed)
        sp = rng.uniform(0.97e5, 1.03e5, (im, jm))
        frac = np.concatenate([[1.0], np.sort(rng.uniform(0.0, 1.0, lm - 1))[::-1], [0.0]])
        if case == "high":
            frac = frac * 0.5  # the floor at half the surface pressure, above the first edge
        phlb = sp[..., None] * frac[None, None, :]
        inp = {
            "temper": rng.uniform(210.0, 300.0, (im, jm, lm)),
            "phlb": phlb,
            "lat": np.linspace(-1.4, 1.4, jm),
            "strat_press": np.sort(rng.uniform(1.0, 700.0, (jm, nstrat)), axis=1)[:, ::-1].copy(),
            "strat_oh": rng.uniform(1e-9, 1e-8, (jm, nstrat)),
            "strat_cl": rng.uniform(1e-10, 1e-9, (jm, nstrat)),
            "strat_od": rng.uniform(1e-11, 1e-10, (jm, nstrat)),
            "oh_3d": rng.uniform(1e5, 1e6, (im, jm, lm)),
        }
        kw = {"oh_scale": 1.1, "cl_scale": 0.9, "o1d_scale": 1.2}
        if case == "trop":
            inp["pl_trop"] = np.sort(rng.uniform(2e4, 1.05e5, (im, jm, ntrop)), axis=2)[..., ::-1].copy()
            inp["cl_trop"] = rng.uniform(1e3, 1e4, (im, jm, ntrop))
            inp["soil"] = rng.uniform(0.0, 1.0, (im, jm))
            kw["soil_rate"] = 2e-4
        loss = chem.loss_rates(**inp, **kw)
        for k, v in {**inp, **{k: np.array(v) for k, v in kw.items()}}.items():
            out[f"{case}.{k}"] = np.asarray(v)
        out[f"{case}.loss"] = loss
        out[f"{case}.decay"] = chem.decay(loss, 1800.0)
    return out


def kzz_cases():
    """pydyn.diffusion.calc_kzz on the reference's own synthetic inputs (pydyn.synth.kzz_inputs)
    at two shapes: dkg and the intermediates (ustar, gph, pblh, kvh), with the inputs."""
    out = {}
    for case, (seed, im, jm, lm, lmax) in {"a": (61, 8, 6, 8, 6), "b": (62, 10, 9, 20, 14)}.items():
        args, grid = synth.kzz_inputs(np.random.default_rng(seed), im, jm, lm=lm, lmax_conv=lmax)
        dkg, parts = diffusion.calc_kzz(*args, **grid)
        spm, sp, pu, pv, t, q, surface, at, bt, dxyp, lmax_conv = args
        inp = {"spm": spm, "sp": sp, "pu": pu, "pv": pv, "t": t, "q": q, "at": at, "bt": bt, "dxyp": dxyp,
               "lmax_conv": np.array(lmax_conv), **{f"surface_{k}": v for k, v in surface.items()},
               **{k: np.array(v) for k, v in grid.items()}}  # fmt: skip
        for k, v in {**inp, "dkg": dkg, **parts}.items():
            out[f"{case}.{k}"] = np.asarray(v)
    return out


def main(argv):
    _use_twin()
    if argv[0] == "ops":
        np.savez_compressed(GOLDEN / "ops_tiny.npz", **ops_cases())
    elif argv[0] == "run":
        run_golden(argv[1], int(argv[2]))
    elif argv[0] == "zoom":
        zoom_golden(argv[1], int(argv[2]))
    elif argv[0] == "chem":
        np.savez_compressed(GOLDEN / "chem_loss.npz", **chem_cases())
    elif argv[0] == "kzz":
        np.savez_compressed(GOLDEN / "kzz.npz", **kzz_cases())
    else:
        raise SystemExit(__doc__)


if __name__ == "__main__":
    main(sys.argv[1:])
