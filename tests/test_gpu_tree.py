"""The zoom tree's queued native run (tm5_tree_run) against the reference, bit for bit: intervals
queued in runs of `chunk` as checked steps, each step of a given ndyn captured into a HIP graph
once and replayed, the intervals whose plan fails (Check_CFL cuts) redone from the backup.  Every
interval's (ndyn, n) and, at the end of every chunk (after the host has read the state, so graph
replays alternate with eager work and synchronisation), every array of every region against the
sha1s the reference's zoom path wrote (tests/golden/zoom_*_sha1.json)."""

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


def _compare(host, want, what):
    bad = [k for k in want if "." in k and sha(host[int(k[1])][k.split(".")[1]]) != want[k]]
    assert not bad, f"{what}: differ from the reference: {bad}"


def _queued(name, chunk, limit=None, graphs=True, ensemble=None, mode=None):
    from pydyn_amd import _lib, meteo
    from pydyn_amd.zoom.run import ZoomRun

    golden = f"zoom_{name}_sha1.json" if ensemble is None else f"zoom_{name}e_sha1.json"
    ref = json.loads((GOLDEN / golden).read_text())["intervals"][:limit]
    if ensemble is None:
        pmet, cmet = meteo.zoom(name)
        run = ZoomRun(pmet, cmet, redgrid=meteo.redgrid(name))
    else:
        pmet, cmet = meteo.zoom(name, nt=ensemble[0])
        run = ZoomRun(pmet, cmet, ensemble=meteo.Ensemble(pmet, ensemble[1]))
    if not graphs:
        run.native().option(_lib.TREE_GRAPHS, 0)
    if mode is not None:
        run.native().option(_lib.TREE_MODE, mode)
    i = 0
    while i < len(ref):
        k = min(chunk, len(ref) - i)
        print(f"{name}: intervals {i}..{i + k - 1} queued (graphs {graphs}, mode {mode})", flush=True)
        out = run.run(k)
        for j, got in enumerate(out):
            assert got == (ref[i + j]["ndyn"], ref[i + j]["n"]), f"{name} interval {i + j}: (ndyn, n) {got}"
        i += k
        _compare(run.host(), ref[i - 1], f"{name} after interval {i - 1}")
    return run


@pytest.mark.parametrize(("name", "chunk", "limit"), [
    ("tiny2", 12, None),     # 96 intervals, 14 cut
    ("small2", 8, None),     # 72 intervals, 46 cut
    ("tiny2r", 16, None),    # region 1 on a reduced grid
    ("small2r", 12, None),
    ("c2r", 24, None),       # the production tree on the glb3x2 reduced grid
])  # fmt: skip
def test_queued_tree_matches_reference(dev, name, chunk, limit):
    run = _queued(name, chunk, limit)
    tally = run.native().tally()
    assert tally["kept"] + tally["redone"] == len(run.log)
    assert tally["replayed"] > 0 and tally["captured"] >= 1, tally


def test_queued_production_tree_matches_reference(dev):
    """The production tree (glb3x2 x 34 with a 120 x 70 x 34 child) for all 168 intervals of the
    golden file (7 days, 60 of them cut by Check_CFL), queued in days: the checked steps of both
    step lengths replayed from their graphs and every plan that fails redone."""
    run = _queued("c2", 24)
    tally = run.native().tally()
    assert tally["redone"] >= 1 and tally["replayed"] >= 100 and tally["captured"] >= 2, tally


def test_queued_tree_without_graphs_matches_reference(dev):
    """The same queue with every launch issued (no graph): the reference's bits too."""
    run = _queued("small2", 12, graphs=False)
    assert run.native().tally()["replayed"] == 0


def test_queued_ensemble_tree_matches_reference(dev):
    """4 members x 4 tracers on the tiny tree, queued, against the reference's ensemble run."""
    _queued("tiny2", 8, ensemble=(4, 4))


def test_queued_production_ensemble_tree_matches_reference(dev):
    """2 members x 4 tracers on the production tree, queued in runs of 12 across the cuts of the
    first 40 hours, against the reference's ensemble run (member 2's emission scaled, its
    tracers started from its own seed)."""
    run = _queued("c2", 12, ensemble=(4, 2))
    assert run.native().tally()["redone"] >= 1


@pytest.mark.parametrize("name", ["small2", "c2r"])
def test_queued_tree_rehearsed_plans_match_reference(dev, name):
    """The queue's other mode (the one many tracers take): every interval's planned trial rehearsed
    mass-only and verified on the device, then plain steps with no backup; the reference's bits."""
    from pydyn_amd import _lib

    run = _queued(name, 12, mode=_lib.TREE_REHEARSED)
    tally = run.native().tally()
    assert tally["replayed"] > 0, tally


def test_queued_ensemble_rehearsed_plans_match_reference(dev):
    from pydyn_amd import _lib

    _queued("tiny2", 12, ensemble=(4, 4), mode=_lib.TREE_REHEARSED)


def test_queued_64x4_members_equal_their_plain_queued_runs(dev):
    """The batched form at full size: 64 members x 4 tracers (nt = 256) on the production tree
    for 40 intervals (the last seven cut by Check_CFL), queued in runs of 12 (the mode without a
    backup); every member's tracers and the tree's mass and fluxes against the member's own
    plain queued run (nt = 4, the checked mode with its backup)."""
    import torch

    from pydyn_amd import meteo
    from pydyn_amd.zoom.run import ZoomRun

    pmet, cmet = meteo.zoom("c2", nt=4)
    ens = meteo.Ensemble(pmet, 64)
    batched = ZoomRun(pmet, cmet, ensemble=ens)
    log = []
    for i in range(0, 40, 12):
        print(f"64 x 4: intervals {i}..{min(i + 12, 40) - 1}", flush=True)
        log += batched.run(min(12, 40 - i))
    assert sum(1 for nd, n in log if n > 1) >= 7, log
    for m in range(64):
        plain = ZoomRun(pmet, cmet, ensemble=ens.subset([m]))
        got = []
        for i in range(0, 40, 20):
            got += plain.run(20)
        assert got == log, f"member {m}: (ndyn, n) {got}"
        for r, fields in batched.st.fields.items():
            for k in ("rm", "rxm", "rym", "rzm"):
                a = fields[k].t[..., 4 * m : 4 * m + 4].contiguous()
                b = plain.st.fields[r][k].t.contiguous()
                assert torch.equal(a.view(torch.int64), b.view(torch.int64)), (m, r, k)
            if m in (0, 63):
                for k in ("m", "am", "bm", "cm", "sp", "phlb"):
                    a, b = fields[k].t.contiguous(), plain.st.fields[r][k].t.contiguous()
                    assert torch.equal(a.view(torch.int64), b.view(torch.int64)), (m, r, k)
        plain.close()
        del plain
        if m % 16 == 15:
            print(f"  members 0..{m} equal their plain runs", flush=True)
    batched.close()
