"""The zoom tree on the device against the reference, bit for bit: a global region with one
child (nested synthetic meteo), every array of both regions with its halos, interval by
interval against the sha1s the reference's own zoom path wrote (tests/golden/zoom_*_sha1.json);
where they differ, the oracle's zoom restatement (oracle/zoom.py) names the array and the cell."""

import hashlib
import json
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pydyn_amd import device as D

    D.context().flags(clear=True)
    return D


def sha(a):
    return hashlib.sha1(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()[:20]


def _first_difference(got, oracle_st):
    for r, fields in got.items():
        for k, v in fields.items():
            w = oracle_st[r][k].data
            b1, b2 = np.ascontiguousarray(v).view(np.int64), np.ascontiguousarray(w).view(np.int64)
            if v.shape != w.shape:
                return f"region {r} {k}: shape {v.shape} vs {w.shape}"
            d = b1 != b2
            if d.any():
                idx = tuple(int(x) for x in np.argwhere(d)[0])
                return f"region {r} {k}: {d.sum()} differ, first {idx}: {v[idx]!r} vs oracle {w[idx]!r}"
    return None


def _run(name, limit=None, oracle_every=None, driver="native"):
    from oracle import zoom as OZ
    from pydyn_amd import meteo
    from pydyn_amd.zoom.run import ZoomRun

    ref = json.loads((GOLDEN / f"zoom_{name}_sha1.json").read_text())["intervals"][:limit]
    pmet, cmet = meteo.zoom(name)
    run = ZoomRun(pmet, cmet, redgrid=meteo.redgrid(name), driver=driver)
    orun = OZ.ZoomRun(pmet, cmet, redgrid=meteo.redgrid(name)) if oracle_every else None
    for it, want in enumerate(ref):
        got_nd = run.interval()
        if orun is not None:
            orun.interval()
        assert got_nd == (want["ndyn"], want["n"]), f"{name} interval {it}: (ndyn, n) {got_nd}"
        if oracle_every and it % oracle_every == 0:
            diff = _first_difference(run.host(), orun.st)
            assert diff is None, f"{name} interval {it}: {diff}"
        host = run.host()
        bad = [k for k in want if "." in k and sha(host[int(k[1])][k.split(".")[1]]) != want[k]]
        if bad and orun is not None:
            bad.append(_first_difference(host, orun.st))
        assert not bad, f"{name} interval {it}: differ from the reference: {bad}"
    return run


@pytest.mark.parametrize("driver", ["native", "python"])
def test_zoom_tiny_tree_matches_reference(dev, driver):
    """The library's tree driver (tm5_tree_interval) and the Python restatement over the bodies."""
    _run("tiny2", oracle_every=1, driver=driver)


@pytest.mark.parametrize("driver", ["native", "python"])
def test_zoom_small_tree_matches_reference(dev, driver):
    _run("small2", limit=24, oracle_every=6, driver=driver)


def test_zoom_reduced_grid_tiny_tree_matches_reference(dev):
    """Region 1's x advection on a reduced grid (rows of 1, 3 and 4 merged cells)."""
    _run("tiny2r", oracle_every=1)


def test_zoom_reduced_grid_small_tree_matches_reference(dev):
    _run("small2r", limit=24, oracle_every=6)


def test_zoom_production_tree_matches_oracle_and_reference(dev):
    """glb3x2 x 34 with a 1x1 degree child over 170W-50W, 10N-80N (120 x 70 x 34): the oracle
    for the first interval, the reference's sha1s for as many intervals as the golden file
    holds."""
    _run("c2", limit=1, oracle_every=1)
    _run("c2")


def test_zoom_reduced_grid_production_tree_matches_reference(dev):
    """The production tree with region 1's x advection on the glb3x2 reduced grid (rows 1-7 and
    84-90 merged in clusters of 40, 8, 8, 4, 4, 4, 2 cells)."""
    _run("c2r")


# ------------------------------------------------------------------ ensembles (members x tracers)
def _ensemble(base, ntr, members):
    from pydyn_amd import meteo

    pmet, cmet = meteo.zoom(base, nt=ntr)
    return pmet, cmet, meteo.Ensemble(pmet, members)


def test_zoom_ensemble_tiny_matches_reference(dev):
    """4 members x 4 tracers on the tiny tree: every array against the reference's run of the
    same ensemble (each member's emission scaled into its first tracer)."""
    from oracle import zoom as OZ
    from pydyn_amd.zoom.run import ZoomRun

    pmet, cmet, ens = _ensemble("tiny2", 4, 4)
    ref = json.loads((GOLDEN / "zoom_tiny2e_sha1.json").read_text())["intervals"]
    run = ZoomRun(pmet, cmet, ensemble=ens)
    orun = OZ.ZoomRun(pmet, cmet, ensemble=ens)
    for it, want in enumerate(ref):
        assert run.interval() == orun.interval() == (want["ndyn"], want["n"]), f"interval {it}"
        host = run.host()
        bad = [k for k in want if "." in k and sha(host[int(k[1])][k.split(".")[1]]) != want[k]]
        if bad:
            bad.append(_first_difference(host, orun.st))
        assert not bad, f"tiny2e interval {it}: {bad}"


def _member_slices_equal(batched, plain, m, ntr):
    import torch

    for r, fields in batched.st.fields.items():
        for k in ("rm", "rxm", "rym", "rzm"):
            a = fields[k].t[..., m * ntr : (m + 1) * ntr]
            b = plain.st.fields[r][k].t
            assert torch.equal(a.contiguous().view(torch.int64), b.contiguous().view(torch.int64)), (m, r, k)
        for k in ("m", "am", "bm", "cm", "sp", "phlb"):
            a, b = fields[k].t, plain.st.fields[r][k].t
            assert torch.equal(a.contiguous().view(torch.int64), b.contiguous().view(torch.int64)), (m, r, k)


@pytest.mark.parametrize(("base", "ntr", "members", "intervals", "check"), [
    ("tiny2", 4, 4, 48, (0, 1, 2, 3)),
    ("tiny2", 4, 8, 48, (0, 5, 7)),
    ("c2", 4, 64, 2, (0, 63)),
])  # fmt: skip
def test_zoom_ensemble_members_equal_their_plain_runs(dev, base, ntr, members, intervals, check):
    """The batched form (all members' tracers side by side in one run) gives every member the bits
    of its own plain run; at c2 the 64 members x 4 tracers of the production tree."""
    from pydyn_amd.zoom.run import ZoomRun

    pmet, cmet, ens = _ensemble(base, ntr, members)
    batched = ZoomRun(pmet, cmet, ensemble=ens)
    plains = {m: ZoomRun(pmet, cmet, ensemble=ens.subset([m])) for m in check}
    for it in range(intervals):
        nd = batched.interval()
        for m, run in plains.items():
            assert run.interval() == nd, (it, m)
    for m, run in plains.items():
        _member_slices_equal(batched, run, m, ntr)


@pytest.mark.parametrize("shape", [
    {"y_many": 3, "range_min": 5, "apply": 8},
    {"y_many": 16, "range_min": 1, "apply": 32},
    {"y_many": 8, "range_min": 64, "apply": 16},
])  # fmt: skip
def test_many_tracer_tile_shapes_change_no_bits(dev, shape):
    """The many-tracer knobs of tm5_ctx_set_tile (y tile, fewest tracers a range, apply tile)
    split the same operations differently: 16 members x 4 tracers (nt = 64, tracer ranges in
    every sweep and the split column form) on the tiny tree with shapes unlike the tuned ones,
    every checked member against its plain run (nt = 4, which these knobs do not shape)."""
    from pydyn_amd.zoom.run import ZoomRun

    D = dev
    ctx = D.context()
    saved = {k: ctx.tile(k) for k in ("y_many", "range_min", "apply")}
    try:
        for k, v in shape.items():
            ctx.set_tile(k, v)
        pmet, cmet, ens = _ensemble("tiny2", 4, 16)
        batched = ZoomRun(pmet, cmet, ensemble=ens)
        plains = {m: ZoomRun(pmet, cmet, ensemble=ens.subset([m])) for m in (0, 9, 15)}
        for it in range(24):
            nd = batched.interval()
            for m, run in plains.items():
                assert run.interval() == nd, (it, m)
        for m, run in plains.items():
            _member_slices_equal(batched, run, m, 4)
    finally:
        for k, v in saved.items():
            ctx.set_tile(k, v)
    with pytest.raises(RuntimeError):
        ctx.set_tile("apply", 12)


# ------------------------------------------------------------------ the physics bodies alone
def test_zoom_physics_bodies_match_oracle(dev):
    """convec_region, sources_region, chemistry_region and assign of pydyn_amd.zoom.lib (the
    body names of pydyn/zoom/host.py:42-48) on both regions of the tiny tree, against the
    oracle's bodies on the same state."""
    import torch

    from oracle import zoom as OZ
    from pydyn_amd import meteo
    from pydyn_amd.zoom import lib as L
    from pydyn_amd.zoom.run import ZoomRun

    D = dev
    pmet, cmet = meteo.zoom("tiny2")
    run, orun = ZoomRun(pmet, cmet), OZ.ZoomRun(pmet, cmet)
    run.interval(), orun.interval()  # a state with slopes
    assert _first_difference(run.host(), orun.st) is None
    rng = np.random.default_rng(9)
    tree, otree = run.zoom, orun.tree
    for r in (1, 2):
        g = tree.g(r)
        conv = {k: D.upload(np.asarray(v)) for k, v in orun.conv[r].items() if k not in ("cloud_top", "cloud_lfs", "lmax")}
        for k in ("cloud_top", "cloud_lfs"):
            conv[k] = D.upload(np.asarray(orun.conv[r][k]).astype(np.int32)).to(torch.int32)
        conv["lmax"] = orun.conv[r]["lmax"]
        L.convec_region(tree, run.st, r, conv, 1800)
        OZ.convec_region(otree, orun.st, r, orun.conv[r], 1800)
        assert _first_difference(run.host(), orun.st) is None, f"convec_region {r}"
        x = rng.uniform(0.0, 1e-3, (g.im, g.jm)) * np.abs(orun.st[r]["rm"].data).max()
        L.sources_region(tree, run.st, r, [(D.upload(x), 1, 1)])
        OZ.sources_region(otree, orun.st, r, [(x, 1, 1)])
        assert _first_difference(run.host(), orun.st) is None, f"sources_region {r}"
        sc = OZ.scope(otree, r)
        shape = (sc[0].stop - sc[0].start, sc[1].stop - sc[1].start, tree.lm)
        decay = rng.uniform(0.5, 1.0, shape)
        L.chemistry_region(tree, run.st, r, D.upload(decay), [1])
        OZ.chemistry_region(otree, orun.st, r, decay, [1])
        assert _first_difference(run.host(), orun.st) is None, f"chemistry_region {r}"
    # assign: region 2's tracers from a second run's
    run2 = ZoomRun(pmet, cmet)
    L.assign(run.st, run2.st, 2, ("rm", "rzm"))
    got, want = run.host(), run2.host()
    for k in ("rm", "rzm"):
        assert np.array_equal(got[2][k].view(np.int64), want[2][k].view(np.int64)), k
