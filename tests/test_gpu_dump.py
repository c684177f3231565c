"""The model's dump format as the way into and out of the device state (pydyn_amd.dump): the
oracle's tiny-tree state after two intervals written as dump events, one per region, read into a
fresh device run (Event.device: the file's Fortran-order bytes uploaded as they lie), which then
runs the next intervals to the oracle's bits; and the device state dumped back out reads as the
oracle's state, array by array."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pydyn_amd import device as D

    D.context().flags(clear=True)
    return D


def _write_state(dump, path, seq, st, zoom_or_tree, lm, nt):
    for r in (1, 2):
        g = zoom_or_tree.g(r)
        meta = {"im": g.im, "jm": g.jm, "lm": lm, "ntracet": nt}
        arrays = {k: (np.asarray(st[r][k].data), st[r][k].lower) for k in ("m", "rm", "rxm", "rym", "rzm", "am", "bm",
                                                                            "cm", "pu", "pv", "sp", "phlb", "dxyp")}  # fmt: skip
        dump.write_event(path, seq, "state", r, meta, arrays)


def test_a_device_run_resumes_from_a_dump(dev, tmp_path):
    from oracle import zoom as OZ
    from pydyn_amd import dump, meteo
    from pydyn_amd.zoom.run import ZoomRun

    pmet, cmet = meteo.zoom("tiny2")
    orun = OZ.ZoomRun(pmet, cmet)
    for _ in range(2):
        orun.interval()
    _write_state(dump, tmp_path / "in", 1, orun.st, orun.tree, pmet.g.lm, pmet.nt)

    run = ZoomRun(pmet, cmet)
    events = dump.load_events(tmp_path / "in", tags=("state",))
    assert [e.region for e in events] == [1, 2]
    for ev in events:
        dump.to_state(ev, run.st)
    run.t = orun.t
    for it in range(3):
        assert run.interval() == orun.interval(), it
        got = run.host()
        for r in (1, 2):
            for k, v in got[r].items():
                w = np.asarray(orun.st[r][k].data)
                assert np.array_equal(v.view(np.int64), w.view(np.int64)), (it, r, k)

    # the device state dumped (host arrays with the declared bounds) reads back as the oracle's
    host = run.host()
    for r in (1, 2):
        g = run.zoom.g(r)
        arrays = {k: (host[r][k], run.st[r][k].lower) for k in ("m", "rm", "rzm", "am", "cm")}
        dump.write_event(tmp_path / "out", 9, "state", r, {"im": g.im, "jm": g.jm, "lm": run.zoom.lm}, arrays)
    for ev in dump.load_events(tmp_path / "out"):
        for k in ev.arrays:
            assert np.array_equal(ev[k].data, np.asarray(orun.st[ev.region][k].data)), (ev.region, k)
        assert np.array_equal(ev.interior("m"), np.asarray(orun.st[ev.region]["m"].data)[2:-2, 2:-2, :])
        # the device upload of a dumped array is the array
        assert np.array_equal(dev.download(ev.device("rm")), ev["rm"].data)


def test_to_state_rejects_other_bounds(dev, tmp_path):
    from pydyn_amd import dump, meteo
    from pydyn_amd.zoom.run import ZoomRun

    pmet, cmet = meteo.zoom("tiny2")
    run = ZoomRun(pmet, cmet)
    g = run.zoom.g(1)
    path = dump.write_event(tmp_path, 1, "state", 1, {"im": g.im, "jm": g.jm, "lm": run.zoom.lm},
                            {"m": (np.ones((g.im, g.jm, run.zoom.lm)), (1, 1, 1))})  # fmt: skip
    with pytest.raises(dump.DumpError, match="declares"):
        dump.to_state(dump.Event(path), run.st)
