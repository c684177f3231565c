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


def _run(name, limit=None, oracle_every=None):
    from oracle import zoom as OZ
    from pydyn_amd import meteo
    from pydyn_amd.zoom.run import ZoomRun

    ref = json.loads((GOLDEN / f"zoom_{name}_sha1.json").read_text())["intervals"][:limit]
    pmet, cmet = meteo.zoom(name)
    run = ZoomRun(pmet, cmet, redgrid=meteo.redgrid(name))
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


def test_zoom_tiny_tree_matches_reference(dev):
    _run("tiny2", oracle_every=1)


def test_zoom_small_tree_matches_reference(dev):
    _run("small2", limit=24, oracle_every=6)


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
