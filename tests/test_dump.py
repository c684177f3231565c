"""The dump event format (pydyn_amd.dump): the reference's known answers for Fortran-bounds
indexing and malformed headers (tests/test_dump.py of the reference), and write -> read round
trips of every kind of header value and array."""

import numpy as np
import pytest

from pydyn_amd import dump


def test_farray_uses_fortran_bounds():
    """The reference's own case (tests/test_dump.py:11-21)."""
    data = np.arange(4 * 3, dtype=float).reshape(4, 3, order="F")
    a = dump.FArray(data, lower=(-1, 0))
    assert a.upper == (2, 2)
    assert a[-1, 0] == data[0, 0]
    assert a[2, 2] == data[3, 2]
    np.testing.assert_array_equal(a[0:1, 1:2], data[1:3, 1:3])
    np.testing.assert_array_equal(a[0:1], data[1:3])  # trailing axes whole
    assert a[0:1, 2].shape == (2,)
    with pytest.raises(IndexError):
        a[-1:2:2]


@pytest.mark.parametrize("text, why", [
    ("seq = 1\nregion = 1\nim = 2\njm = 2\nlm = 1\n", "no tag"),
    ("tag = meteo\nseq = 1\nregion = 1\nim = 2\njm = 2\nlm = 1\narray m = 1 2 1\n", "not an array line"),
    ("tag = meteo\nseq = 1\nregion = 1\nim = 2\njm = 2\nlm = 1\narray m = 5 1\n", "bad array bounds"),
    ("tag = meteo\nseq = 1\nregion = 1\nim = 2\njm = 2\nlm = 1\narray ../m = 1 2\n", "or name"),
])  # fmt: skip
def test_a_malformed_header_is_a_dump_error(tmp_path, text, why):
    """The reference's malformed headers (tests/test_dump.py:24-40) name their fault."""
    path = tmp_path / "e00001_meteo_r1_p000.meta"
    path.write_text(text)
    with pytest.raises(dump.DumpError, match=why):
        dump.Event(path)


def test_round_trip_of_an_event(tmp_path):
    rng = np.random.default_rng(7)
    im, jm, lm, nt = 5, 4, 3, 2
    m = rng.uniform(1.0, 2.0, (im + 4, jm + 4, lm))
    rm = rng.standard_normal((im + 4, jm + 4, lm, nt))
    am = rng.standard_normal((im + 2, jm + 2, lm + 2))
    flags = rng.integers(-3, 9, (im, jm)).astype(np.int32)
    meta = {"im": im, "jm": jm, "lm": lm, "ntracet": nt, "dt": 1800.5, "cyclic": True, "sizes": (3, 4, 5),
            "label": "two words"}  # fmt: skip
    path = dump.write_event(tmp_path, 12, "advect", 1, meta,
                            {"m": (m, (-1, -1, 1)), "rm": (rm, (-1, -1, 1, 1)), "am": (am, (0, 0, 0)),
                             "flags": (flags, (1, 1))})  # fmt: skip
    assert path.name == "e00012_advect_r1_p000.meta"
    ev = dump.Event(path)
    assert (ev.tag, ev.seq, ev.region, ev.rank) == ("advect", 12, 1, 0)
    assert ev.meta["dt"] == 1800.5 and ev.meta["cyclic"] is True and ev.meta["sizes"] == (3, 4, 5)
    assert ev.meta["label"] == "two words" and ev.meta["ntracet"] == nt
    assert ev.bounds["rm"] == [(-1, im + 2), (-1, jm + 2), (1, lm), (1, nt)]
    assert ev.dtypes["flags"] == np.dtype("<i4")
    for name, want in (("m", m), ("rm", rm), ("am", am), ("flags", flags)):
        assert np.array_equal(ev[name].data, want) and ev[name].data.dtype == want.dtype, name
    # the interior sections of the reference's layout
    assert np.array_equal(ev.interior("m"), m[2 : im + 2, 2 : jm + 2, :])
    assert np.array_equal(ev.interior("am"), am[0 : im + 1, 1 : jm + 1, 1 : lm + 1])
    assert np.array_equal(ev["rm"][1, 1, 1], rm[2, 2, 0])
    # the .bin is Fortran order: i fastest
    raw = np.fromfile(f"{ev.prefix}_m.bin", dtype="<f8")
    assert np.array_equal(raw[: im + 4], m[:, 0, 0])


def test_a_short_bin_file_is_named(tmp_path):
    path = dump.write_event(tmp_path, 1, "meteo", 2, {"im": 2, "jm": 2, "lm": 1}, {"m": (np.ones((2, 2, 1)), (1, 1, 1))})
    ev = dump.Event(path)
    np.ones(3).tofile(f"{ev.prefix}_m.bin")
    with pytest.raises(dump.DumpError, match="3 values"):
        ev["m"]


def test_load_events_orders_and_filters(tmp_path):
    head = {"im": 2, "jm": 2, "lm": 1}
    a = {"m": (np.ones((2, 2, 1)), (1, 1, 1))}
    for seq, tag, region, rank in ((3, "meteo", 1, 0), (1, "advect", 2, 0), (2, "advect", 1, 1), (2, "advect", 1, 0)):
        dump.write_event(tmp_path, seq, tag, region, head, a, rank=rank)
    (tmp_path / "notes.txt").write_text("not an event")
    evs = dump.load_events(tmp_path)
    assert [(e.seq, e.rank) for e in evs] == [(1, 0), (2, 0), (2, 1), (3, 0)]
    assert [e.seq for e in dump.load_events(tmp_path, tags=("advect",), region=1)] == [2, 2]
    assert [e.rank for e in dump.load_events(tmp_path, rank=1)] == [1]
