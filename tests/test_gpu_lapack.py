"""The LAPACK-kind inverse on the device (k_lapack_inverse: DGETRF2 + DGETRI per matrix, two
matrices a wave) and convec with it, bit for bit against the oracle (oracle/lapack.c, itself pinned to
the reference's dgetrf_ / dgetri_ twin by tests/test_oracle_lapack.py)."""

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
    from pydyn_amd.gpu import kernels as K

    D.context().flags(clear=True)
    return D, K


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


@pytest.mark.parametrize("n", [4, 5, 8, 10, 15, 20, 25, 26])
def test_inverse_rows_matches_oracle(dev, n):
    import torch

    from oracle import lapack

    D, K = dev
    rng = np.random.default_rng(n)
    a = rng.standard_normal((203, n, n))  # pivoting throughout; a partial last workgroup
    a[:50] = np.eye(n) + 0.1 * a[:50]
    a[50:60] *= 1e-300  # pivots below DLAMCH('S'): the division branch
    keep = (rng.random((40, n, n)) < 0.5) | np.eye(n, dtype=bool)  # a nonzero diagonal: nonsingular
    a[61:140:2] *= keep  # exact zeros (the skips) in every other matrix: the halves of a wave diverge
    y, info = K.inverse_rows(torch.as_tensor(a, device="cuda"))
    assert int(info.cpu()[0]) == 0
    assert np.array_equal(_bits(y.cpu().numpy()), _bits(lapack.inverse_rows(a)))


def test_inverse_rows_reports_a_zero_pivot(dev):
    import torch

    D, K = dev
    a = np.stack([np.eye(5), np.ones((5, 5)), 2.0 * np.eye(5)])  # the second is singular
    y, info = K.inverse_rows(torch.as_tensor(a, device="cuda"))
    assert int(info.cpu()[0]) == 1
    y = y.cpu().numpy()
    assert np.array_equal(y[0], np.eye(5)) and np.array_equal(y[2], 0.5 * np.eye(5))


@pytest.mark.parametrize("masked", [False, True])
def test_convec_lapack_matches_oracle(dev, masked):
    """convec(..., inverse='lapack') on 37 x 7 columns of glb3x2 (34 levels over 26) with three
    tracers, on all columns or a column list."""
    import torch

    from oracle import convection, physics
    from pydyn_amd import meteo

    D, K = dev
    met = meteo.Meteo(meteo.grid("glb3x2"), seed=6)
    sl = (slice(0, 37), slice(20, 27))
    lmax, nt = met.g.lmax, 3
    m = np.ascontiguousarray(met.m[sl])
    im, jm, lm = m.shape
    dkg = np.ascontiguousarray(physics.dkg(met.kvh[sl], m, met.gph[sl], lmax))
    c = {k: np.ascontiguousarray(getattr(met, k)[sl]) for k in ("entu", "detu", "entd", "detd", "cloud_top", "cloud_lfs")}
    rng = np.random.default_rng(21)
    rm = rng.uniform(0.5, 1.5, (im, jm, lm, nt)) * m[..., None] * 1.8e-6
    st = {"m": m, "rm": rm, **{k: rm * rng.uniform(-1.0, 1.0, rm.shape) for k in ("rxm", "rym", "rzm")}}
    conv = {k: D.upload(c[k]) for k in ("entu", "detu", "entd", "detd")}
    conv["dkg"] = D.upload(dkg)
    for k in ("cloud_top", "cloud_lfs"):
        conv[k] = D.upload(c[k].astype(np.int32)).to(torch.int32)
    cols, listed = None, np.ones((im, jm), dtype=bool)
    if masked:
        flat = np.sort(rng.choice(im * jm, size=im * jm // 2, replace=False)).astype(np.int32)
        cols = torch.as_tensor(flat, device="cuda")
        listed[...] = False
        listed[flat // jm, flat % jm] = True
    dst = {k: D.upload(st[k]) for k in FIELDS}
    K.convec(*(dst[k] for k in FIELDS), conv, 1800.0, lmax, cols=cols, inverse="lapack")
    full = {k: st[k].copy() for k in FIELDS}
    convection.convec(*(full[k] for k in FIELDS), c["entu"], c["detu"], c["entd"], c["detd"], c["cloud_top"],
                      c["cloud_lfs"], dkg, 1800.0, lmax, inverse="lapack")  # fmt: skip
    crout = {k: st[k].copy() for k in FIELDS}
    convection.convec(*(crout[k] for k in FIELDS), c["entu"], c["detu"], c["entd"], c["detd"], c["cloud_top"],
                      c["cloud_lfs"], dkg, 1800.0, lmax)  # fmt: skip
    differs = False
    for k in FIELDS[1:]:
        got = D.download(dst[k])
        assert np.array_equal(_bits(got[listed]), _bits(full[k][listed])), f"{k} listed"
        assert np.array_equal(_bits(got[~listed]), _bits(st[k][~listed])), f"{k} not listed"
        differs = differs or not np.array_equal(_bits(full[k]), _bits(crout[k]))
    assert differs  # the two kinds are not the same bits, so the test tells them apart


def test_convec_rejects_other_inverses(dev):
    D, K = dev
    with pytest.raises(ValueError, match="'crout' or 'lapack'"):
        K.convec(None, None, None, None, None, {}, 1.0, 4, inverse="numpy")
