"""The rest of the zoom seam a reference caller uses (pydyn/gpu/zoom.py:38-101, 917-1041): a
device state made by zeros, moved by upload / download, and intervals submitted through
Pipeline -- the production tree's 168 intervals queued as the reference's graph-replayed path is,
and a Python-driven run submitted call by call -- against the reference's sha1s."""

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


def _golden(name):
    return json.loads((GOLDEN / f"zoom_{name}_sha1.json").read_text())["intervals"]


def _compare(host, want, what):
    bad = [k for k in want if "." in k and sha(host[int(k[1])][k.split(".")[1]]) != want[k]]
    assert not bad, f"{what}: differ from the reference: {bad}"


def test_zeros_upload_download_round_trip(dev):
    """zeros gives every declared array at zero; download then upload gives back the same bits,
    bounds and shapes, and a run continued from the uploaded copy stays on the reference."""
    from pydyn_amd import meteo
    from pydyn_amd.zoom import lib as L
    from pydyn_amd.zoom.run import ZoomRun

    pmet, cmet = meteo.zoom("tiny2")
    run = ZoomRun(pmet, cmet)
    z = L.zeros(run.zoom, 1)
    for r, f in L.download(z).fields.items():
        for k, v in f.items():
            assert v.lower == run.st[r][k].lower and not v.data.any(), (r, k)
    for _ in range(3):
        run.interval()
    host = L.download(run.st)
    back = L.download(L.upload(host))
    for r, f in host.fields.items():
        for k, v in f.items():
            w = back.fields[r][k]
            assert w.lower == v.lower and np.array_equal(w.data.view(np.int64), v.data.view(np.int64)), (r, k)
    ref = _golden("tiny2")
    _compare({r: {k: v.data for k, v in f.items()} for r, f in host.fields.items()}, ref[2], "tiny2 after 3")
    with pytest.raises(ValueError, match="Zoom"):
        L.upload(type("S", (), {"fields": host.fields})())


@pytest.mark.parametrize(("name", "depth", "repeat"), [("c2", 24, "interval"), ("small2", 8, None)])
def test_pipeline_queues_tree_intervals(dev, name, depth, repeat):
    """Pipeline.submit(run.interval, repeat=...) for every interval of the golden file: the (ndyn, n)
    of each and every array at the end, the reference's; with repeat, steps replayed from graphs."""
    from pydyn_amd import meteo
    from pydyn_amd.zoom import lib as L
    from pydyn_amd.zoom.run import ZoomRun

    ref = _golden(name)
    pmet, cmet = meteo.zoom(name)
    run = ZoomRun(pmet, cmet)
    pipe = L.Pipeline(run.zoom, run.st, depth)
    got = []
    for k in range(len(ref)):
        pipe.submit(run.interval, repeat=repeat)
        if (k + 1) % 48 == 0:
            print(f"{name}: {k + 1} intervals submitted", flush=True)
            got.extend(pipe.drain())
    got.extend(pipe.drain())
    assert got == [(x["ndyn"], x["n"]) for x in ref]
    _compare(run.host(), ref[-1], f"{name} after {len(ref)} intervals")
    assert pipe.tally["kept"] + pipe.tally["redone"] == len(ref), pipe.tally
    assert (pipe.tally["replayed"] > 0) == (repeat is not None), pipe.tally


def test_pipeline_runs_other_calls_in_order(dev):
    """Calls that are not a native interval (here the Python driver's intervals) run as submitted,
    in order, their results returned by drain."""
    from pydyn_amd import meteo
    from pydyn_amd.zoom import lib as L
    from pydyn_amd.zoom.run import ZoomRun

    ref = _golden("tiny2")[:12]
    pmet, cmet = meteo.zoom("tiny2")
    run = ZoomRun(pmet, cmet, driver="python")
    pipe = L.Pipeline(run.zoom, run.st, 2)
    for _ in ref:
        pipe.submit(run.interval)
    assert pipe.drain() == [(x["ndyn"], x["n"]) for x in ref]
    _compare(run.host(), ref[-1], "tiny2 after 12")
