"""tm5_convec_noalloc, the allocation-free convec, against tm5_convec through the C ABI: the same
bits on two calls in a row (the second on a smaller grid, so the static buffer holds the first
call's matrices when the second starts), in bands of rows and in one band; and the rejection of a
row of columns larger than the static buffer."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("m", "rm", "rxm", "rym", "rzm")


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pydyn_amd import device as D

    D.context().flags(clear=True)
    return D


def _case(sl, nt, seed):
    """A convec case on the glb3x2 meteo's columns sl: inputs in the pydyn layout."""
    from oracle import physics
    from pydyn_amd import meteo

    met = meteo.Meteo(meteo.grid("glb3x2"), seed=4)
    m = np.ascontiguousarray(met.m[sl])
    lmax = met.g.lmax
    dkg = np.ascontiguousarray(physics.dkg(met.kvh[sl], m, met.gph[sl], lmax))
    c = {k: np.ascontiguousarray(getattr(met, k)[sl]) for k in ("entu", "detu", "entd", "detd", "cloud_top", "cloud_lfs")}
    rng = np.random.default_rng(seed)
    rm = rng.uniform(0.5, 1.5, (*m.shape, nt)) * m[..., None] * 1.8e-6
    st = {"m": m, "rm": rm, **{k: rm * rng.uniform(-1.0, 1.0, rm.shape) for k in ("rxm", "rym", "rzm")}}
    return st, c, dkg, lmax


def _run(D, entry, st, c, dkg, lmax, dt=1800.0):
    """One call of `entry` (tm5_convec or tm5_convec_noalloc) on fresh device copies of st; the
    tracers after it, on the host."""
    import torch

    from pydyn_amd import _lib

    dst = {k: D.upload(st[k]) for k in FIELDS}
    conv = {k: D.upload(c[k]) for k in ("entu", "detu", "entd", "detd")}
    conv["dkg"] = D.upload(dkg)
    ci = {k: D.upload(c[k].astype(np.int32)).to(torch.int32) for k in ("cloud_top", "cloud_lfs")}
    im, jm, lm = st["m"].shape
    tr = D.tracers(*(dst[k] for k in ("rm", "rxm", "rym", "rzm")))
    ctx = D.context()
    print(f"{entry}: {im} x {jm} x {lm}, nt {tr.nt}", flush=True)
    rc = getattr(_lib.lib(), entry)(
        ctx.ptr, im, jm, lm, D.view(dst["m"]), ctypes.byref(tr),
        *(D.view(conv[k]) for k in ("entu", "detu", "entd", "detd")),
        D.iview(ci["cloud_top"]), D.iview(ci["cloud_lfs"]), D.view(conv["dkg"]), dt, lmax)  # fmt: skip
    _lib.check(rc, entry)
    ctx.raise_errors()
    return {k: D.download(dst[k]) for k in FIELDS[1:]}


def _assert_bitwise(got, want, what):
    differ = np.asarray(got).view(np.int64) != np.asarray(want).view(np.int64)
    assert not differ.any(), f"{what}: {differ.sum()} of {differ.size} differ"


def test_noalloc_convec_equals_convec_twice_smaller_second(dev):
    """Call 1: the whole glb3x2 grid (120 x 90 columns, lmax 26, 12 tracers), two bands of rows
    (51 rows fill the 4 Mi-double buffer); call 2: a 37 x 7 patch with 9 tracers in one band."""
    from pydyn_amd import _lib

    D = dev
    full = (slice(None), slice(None))
    st, c, dkg, lmax = _case(full, 12, 21)
    im = st["m"].shape[0]
    assert _lib.NOALLOC_DOUBLES // (im * lmax * lmax) < st["m"].shape[1]  # more than one band
    want = _run(D, "tm5_convec", st, c, dkg, lmax)
    got = _run(D, "tm5_convec_noalloc", st, c, dkg, lmax)
    for k in want:
        _assert_bitwise(got[k], want[k], f"call 1 {k}")
        assert not np.array_equal(want[k], st[k]), f"call 1 {k}: convec changed nothing"
    patch = (slice(0, 37), slice(40, 47))
    st2, c2, dkg2, _ = _case(patch, 9, 22)
    want2 = _run(D, "tm5_convec", st2, c2, dkg2, lmax)
    got2 = _run(D, "tm5_convec_noalloc", st2, c2, dkg2, lmax)
    for k in want2:
        _assert_bitwise(got2[k], want2[k], f"call 2 {k}")


def test_noalloc_convec_rejects_a_row_beyond_its_buffer(dev):
    """im lmax^2 above the buffer: ValueError before anything is queued (the arrays are real and
    of the stated shape all the same)."""
    import torch

    from pydyn_amd import _lib

    D = dev
    lmax, lm, nt = 26, 26, 8
    im = _lib.NOALLOC_DOUBLES // (lmax * lmax) + 1
    m = D.zeros((im, 1, lm))
    q = [D.zeros((im, 1, lm, nt)) for _ in range(4)]
    e = D.zeros((im, 1, lm))
    top = torch.zeros((1, im), dtype=torch.int32, device="cuda")
    tr = D.tracers(*q)
    iv = _lib.IView(top.data_ptr(), (ctypes.c_int64 * 2)(1, im))
    rc = _lib.lib().tm5_convec_noalloc(D.context().ptr, im, 1, lm, D.view(m), ctypes.byref(tr), *([D.view(e)] * 4),
                                       iv, iv, D.view(e), 1800.0, lmax)  # fmt: skip
    assert rc == _lib.ESCRATCH
    with pytest.raises(ValueError, match="static buffer"):
        _lib.check(rc, "tm5_convec_noalloc")
