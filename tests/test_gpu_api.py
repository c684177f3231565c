"""The operator layer's names beyond what the step itself calls (pydyn/gpu/kernels.py): reduced
rows and the per-row Courant maximum of advect_x_rows / mass_x_rows, bare z sub-steps, cluster
sums, convection on a column list, and dynamv on a slice of region 1's halo'd bm.  Each against
the oracle on the same inputs, bit for bit.
"""

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


def assert_bitwise(got, want, what=""):
    got, want = np.asarray(got, dtype=float), np.asarray(want, dtype=float)
    assert got.shape == want.shape, what
    differ = got.view(np.int64) != want.view(np.int64)
    assert not differ.any(), f"{what}: {differ.sum()} of {differ.size} differ"


def records(op, seed=3):
    """The inputs (m, flux, tracers) of every `op` operator of one hourly interval of the small
    grid, as the oracle's run recorded them, with slopes the limiter clips."""
    from oracle import adjoint as A
    from pydyn_amd import meteo

    met = meteo.Meteo(meteo.grid("small"), seed=seed)
    rm = met.state()["rm"]
    rng = np.random.default_rng(5)
    init = [rm] + [2.0 * rm * rng.uniform(-1.0, 1.0, rm.shape) for _ in range(3)]
    W = A.Window(met, 1, init=init)
    return [(m.copy(), flux.copy(), [x.copy() for x in Q]) for o, _, m, flux, Q in W.ops if o == op]


def test_advect_x_rows_whole_rows(dev):
    from oracle import advection

    D, K = dev
    m, a, q = records("x")[0]
    a = 3.0 * a  # rows sub-cycle
    im, jm, lm = m.shape
    m0, a0 = m.copy(), a.copy()
    dm, dq, da = D.upload(m), [D.upload(x) for x in q], D.upload(a)
    nloop, amax = K.advect_x_rows(dm, *dq, da, None)
    want_n = advection.advect_x(m, *q, a, True)  # in place; (jm, lm)
    for k, (g, w) in enumerate(zip([dm, *dq, da], [m, *q, a])):
        assert_bitwise(D.download(g), w, FIELDS[k] if k < 5 else "am")
    got_n = nloop.cpu().numpy().reshape(lm, jm).T  # row j + jm l
    assert np.array_equal(got_n, want_n)
    assert want_n.max() > 1 and want_n.min() == 1  # some rows sub-cycle, some not
    # rows that did not sub-cycle: the largest |Courant| of their one trial sub-step
    up_lo = np.concatenate([m0[-1:], m0], axis=0)  # upwind cell of face k: k-1 (periodic) or k
    up_hi = np.concatenate([m0, m0[:1]], axis=0)
    c = np.abs(a0 / np.where(a0 >= 0.0, up_lo, up_hi))
    got_a = amax.cpu().numpy().reshape(lm, jm).T
    one = want_n == 1
    assert np.array_equal(got_a[one], c.max(axis=0)[one])
    assert np.all((got_a > 0.0) & (got_a < 1.0))


@pytest.mark.parametrize("mass_only", [False, True])
def test_x_rows_reduced(dev, mass_only):
    """Rows of nrow[j] cells (a reduced latitude's clusters first): each the periodic row of its
    first nrow[j] cells and faces 0..nrow[j]; the cells beyond untouched."""
    from oracle import advection, cfl

    D, K = dev
    torch = D._torch()
    m, a, q = records("x")[0]
    a = 2.5 * a  # every row sub-cycles
    im, jm, lm = m.shape
    nrow = np.array([im if j in (0, jm - 1) else (im // 2, im // 3, im // 4)[j % 3] for j in range(jm)],
                    dtype=np.int32)  # fmt: skip
    for j, n in enumerate(nrow):
        a[n, j] = a[0, j]  # a reduced row closes at its own last face
    dm, dq, da = D.upload(m), [D.upload(x) for x in q], D.upload(a)
    rows = torch.as_tensor(nrow, device="cuda")
    if mass_only:
        ok = K.mass_x_rows(dm, da, rows, max_nloop=10)
        assert ok
    else:
        nloop, _ = K.advect_x_rows(dm, *dq, da, rows)
    for j, n in enumerate(nrow):
        sm, sa = m[:n, j : j + 1].copy(), a[: n + 1, j : j + 1].copy()
        sq = [x[:n, j : j + 1].copy() for x in q]
        if mass_only:
            assert cfl.mass_x(sm, sa, 10)
        else:
            want = advection.advect_x(sm, *sq, sa, True)
            assert np.array_equal(nloop.cpu().numpy().reshape(lm, jm)[:, j], want[0])
            assert want.max() > 1
        assert_bitwise(D.download(dm)[:n, j], sm[:, 0], f"m row {j}")
        assert_bitwise(D.download(dm)[n:, j], m[n:, j], f"m beyond row {j}")
        assert_bitwise(D.download(da)[: n + 1, j], sa[:, 0], f"am row {j}")
        if not mass_only:
            for k in range(4):
                assert_bitwise(D.download(dq[k])[:n, j], sq[k][:, 0], f"{FIELDS[k + 1]} row {j}")
                assert_bitwise(D.download(dq[k])[n:, j], q[k][n:, j], f"{FIELDS[k + 1]} beyond row {j}")


@pytest.mark.parametrize("prelimit", [False, True])
@pytest.mark.parametrize("n_advz", [1, 3])
def test_advect_z_substeps(dev, n_advz, prelimit):
    from oracle import advection
    from oracle.lines import limit

    D, K = dev
    m, cm, q = records("z")[0]
    dm, dq, dc = D.upload(m), [D.upload(x) for x in q], D.upload(cm)
    K.advect_z_substeps(dm, *dq, dc, n_advz, True, prelimit)
    if prelimit:
        q[3][...] = limit(q[3], q[0])
    advection.z_substeps(m, *q, cm, n_advz, True)
    for k, (g, w) in enumerate(zip([dm, *dq, dc], [m, *q, cm])):
        assert_bitwise(D.download(g), w, FIELDS[k] if k < 5 else "cm")


@pytest.mark.parametrize("scale", [1.0, 8.0])
def test_advect_z_many_tracers(dev, scale):
    """dynamw on 40 tracers (tracer ranges over the grid, each range moving the masses through
    the sub-steps on its own, then the masses once), with and without sub-cycling, against the
    oracle."""
    from oracle import advection

    D, K = dev
    m, cm, q = records("z")[0]
    cm = scale * cm
    rng = np.random.default_rng(40)
    q = [np.ascontiguousarray(np.repeat(x, 40, axis=3) * rng.uniform(0.5, 1.5, (*x.shape[:3], 40))) for x in q]
    dm, dq, dc = D.upload(m), [D.upload(x) for x in q], D.upload(cm)
    n = K.advect_z(dm, *dq, dc, True)
    want_n = advection.advect_z(m, *q, cm, True)
    assert n == want_n and (scale == 1.0 or n > 1), (n, want_n)
    for k, (g, w) in enumerate(zip([dm, *dq, dc], [m, *q, cm])):
        assert_bitwise(D.download(g), w, FIELDS[k] if k < 5 else "cm")


@pytest.mark.parametrize("scale", [1.0, 3.0])
def test_advect_x_many_tracers(dev, scale):
    """dynamu on 40 tracers (tracer ranges over the rows, each range searching and moving the row's
    mass on its own, then the masses and am once), with and without sub-cycling rows, against the
    oracle."""
    from oracle import advection

    D, K = dev
    m, a, q = records("x")[0]
    a = scale * a
    rng = np.random.default_rng(41)
    q = [np.ascontiguousarray(np.repeat(x, 40, axis=3) * rng.uniform(0.5, 1.5, (*x.shape[:3], 40))) for x in q]
    dm, dq, da = D.upload(m), [D.upload(x) for x in q], D.upload(a)
    K.advect_x(dm, *dq, da, True)
    rows = advection.advect_x(m, *q, a, True)
    assert scale == 1.0 or int(np.max(rows)) > 1
    for k, (g, w) in enumerate(zip([dm, *dq, da], [m, *q, a])):
        assert_bitwise(D.download(g), w, FIELDS[k] if k < 5 else "am")


@pytest.mark.parametrize("scale", [1.0, 3.0])
def test_advect_y_many_tracers(dev, scale):
    """dynamv on 40 tracers (tracer ranges over the grid, each range moving the line masses
    through the sub-steps in LDS, then the masses and the bm restore once), with and without
    sub-cycling layers, against the oracle."""
    from oracle import advection

    D, K = dev
    m, b, q = records("y")[0]
    b = scale * b
    rng = np.random.default_rng(42)
    q = [np.ascontiguousarray(np.repeat(x, 40, axis=3) * rng.uniform(0.5, 1.5, (*x.shape[:3], 40))) for x in q]
    dm, dq, db = D.upload(m), [D.upload(x) for x in q], D.upload(b)
    K.advect_y(dm, *dq, db)
    layers = advection.advect_y(m, *q, b, True)
    assert scale == 1.0 or int(np.max(layers)) > 1
    for k, (g, w) in enumerate(zip([dm, *dq, db], [m, *q, b])):
        assert_bitwise(D.download(g), w, FIELDS[k] if k < 5 else "bm")


def test_cluster_sum(dev):
    D, K = dev
    rng = np.random.default_rng(4)
    q = rng.standard_normal((30, 4, 3))
    got = D.download(K.cluster_sum(D.upload(q), 5, 6))
    want = q[0:30:6].copy()
    for s in range(1, 6):  # each cluster from its first cell, in index order
        want = want + q[s:30:6]
    assert_bitwise(got, want, "3-d")
    got2 = D.download(K.cluster_sum(D.upload(q[:, :, 0]), 10, 3))
    want2 = (q[0:30:3, :, 0] + q[1:30:3, :, 0]) + q[2:30:3, :, 0]
    assert_bitwise(got2, want2, "2-d")


def test_convec_on_a_column_list(dev):
    """convec's `cols` (flat i * jm + j): the listed columns as the whole-region operator leaves
    them, the others unchanged."""
    import torch

    from oracle import convection
    from pydyn_amd import meteo

    D, K = dev
    met = meteo.Meteo(meteo.grid("small"), seed=5)
    st = met.state()
    g = met.g
    im, jm = g.im, g.jm
    rng = np.random.default_rng(5)
    cols = np.sort(rng.choice(im * jm, size=im * jm // 3, replace=False)).astype(np.int32)
    from oracle import physics

    dkg = physics.dkg(met.kvh, met.m, met.gph, g.lmax)
    conv = {k: D.upload(getattr(met, k)) for k in ("entu", "detu", "entd", "detd")}
    conv["dkg"] = D.upload(dkg)
    for k in ("cloud_top", "cloud_lfs"):
        conv[k] = D.upload(getattr(met, k).astype(np.int32)).to(torch.int32)
    dst = {k: D.upload(st[k]) for k in FIELDS}
    K.convec(*(dst[k] for k in FIELDS), conv, 1800.0, g.lmax, cols=torch.as_tensor(cols, device="cuda"))
    full = {k: st[k].copy() for k in FIELDS}
    convection.convec(*(full[k] for k in FIELDS), met.entu, met.detu, met.entd, met.detd, met.cloud_top,
                      met.cloud_lfs, dkg, 1800.0, g.lmax)  # fmt: skip
    listed = np.zeros((im, jm), dtype=bool)
    listed[cols // jm, cols % jm] = True
    for k in FIELDS[1:]:
        got = D.download(dst[k])
        assert_bitwise(got[listed], full[k][listed], f"{k} listed")
        assert_bitwise(got[~listed], st[k][~listed], f"{k} not listed")


@pytest.mark.parametrize("masked", [False, True])
def test_convec_many_tracers(dev, masked):
    """convec on 12 tracers (the split form: every column's matrix first, then the tracers in
    workgroups of 16 columns) at 34 levels over 26 (levels 32 and 33 above the convective ones),
    on all columns or a column list, against the oracle; and each tracer as one-tracer calls."""
    import torch

    from oracle import convection, physics
    from pydyn_amd import meteo

    D, K = dev
    met = meteo.Meteo(meteo.grid("glb3x2"), seed=4)
    sl = (slice(0, 37), slice(40, 47))  # 37 x 7 columns: a partial workgroup of 16 at each row's end
    lmax, nt = met.g.lmax, 12
    m = np.ascontiguousarray(met.m[sl])
    im, jm, lm = m.shape
    dkg = np.ascontiguousarray(physics.dkg(met.kvh[sl], m, met.gph[sl], lmax))
    c = {k: np.ascontiguousarray(getattr(met, k)[sl]) for k in ("entu", "detu", "entd", "detd", "cloud_top", "cloud_lfs")}
    rng = np.random.default_rng(12)
    rm = rng.uniform(0.5, 1.5, (im, jm, lm, nt)) * m[..., None] * 1.8e-6
    st = {"m": m, "rm": rm, **{k: rm * rng.uniform(-1.0, 1.0, rm.shape) for k in ("rxm", "rym", "rzm")}}
    conv = {k: D.upload(c[k]) for k in ("entu", "detu", "entd", "detd")}
    conv["dkg"] = D.upload(dkg)
    for k in ("cloud_top", "cloud_lfs"):
        conv[k] = D.upload(c[k].astype(np.int32)).to(torch.int32)
    cols = None
    listed = np.ones((im, jm), dtype=bool)
    if masked:
        flat = np.sort(rng.choice(im * jm, size=im * jm // 2, replace=False)).astype(np.int32)
        cols = torch.as_tensor(flat, device="cuda")
        listed[...] = False
        listed[flat // jm, flat % jm] = True
    dst = {k: D.upload(st[k]) for k in FIELDS}
    K.convec(*(dst[k] for k in FIELDS), conv, 1800.0, lmax, cols=cols)
    D.context().raise_errors()
    full = {k: st[k].copy() for k in FIELDS}
    convection.convec(*(full[k] for k in FIELDS), c["entu"], c["detu"], c["entd"], c["detd"], c["cloud_top"],
                      c["cloud_lfs"], dkg, 1800.0, lmax)  # fmt: skip
    for k in FIELDS[1:]:
        got = D.download(dst[k])
        assert_bitwise(got[listed], full[k][listed], f"{k} listed")
        assert_bitwise(got[~listed], st[k][~listed], f"{k} not listed")
    for t in (0, 7, 11):  # one tracer a call: the one-kernel form
        one = {k: D.upload(np.ascontiguousarray(st[k][..., t : t + 1])) for k in FIELDS[1:]}
        K.convec(D.upload(m), *(one[k] for k in FIELDS[1:]), conv, 1800.0, lmax, cols=cols)
        for k in FIELDS[1:]:
            assert_bitwise(D.download(one[k])[..., 0], D.download(dst[k])[..., t], f"{k} tracer {t}")


def test_convec_split_form_full_grid_40_tracers(dev):
    """The split form over the whole 120 x 90 x 34 grid with 40 tracers (tiles of 16 columns over
    every row, the tracers in ranges): every tracer against its own one-tracer call (the one-kernel
    form) and a block of columns against the oracle."""
    import torch

    from oracle import convection, physics
    from pydyn_amd import meteo

    D, K = dev
    met = meteo.Meteo(meteo.grid("glb3x2"), seed=9)
    lmax, nt = met.g.lmax, 40
    m = np.ascontiguousarray(met.m)
    im, jm, lm = m.shape
    dkg = np.ascontiguousarray(physics.dkg(met.kvh, m, met.gph, lmax))
    rng = np.random.default_rng(40)
    rm = rng.uniform(0.5, 1.5, (im, jm, lm, nt)) * m[..., None] * 1.8e-6
    st = {"m": m, "rm": rm, **{k: rm * rng.uniform(-1.0, 1.0, rm.shape) for k in ("rxm", "rym", "rzm")}}
    conv = {k: D.upload(np.ascontiguousarray(getattr(met, k))) for k in ("entu", "detu", "entd", "detd")}
    conv["dkg"] = D.upload(dkg)
    for k in ("cloud_top", "cloud_lfs"):
        conv[k] = D.upload(getattr(met, k).astype(np.int32)).to(torch.int32)
    dst = {k: D.upload(st[k]) for k in FIELDS}
    print(f"split form {im} x {jm} x {lm}, {nt} tracers", flush=True)
    K.convec(*(dst[k] for k in FIELDS), conv, 1800.0, lmax)
    D.context().raise_errors()
    dm = D.upload(m)
    for t in range(nt):
        one = {k: D.upload(np.ascontiguousarray(st[k][..., t : t + 1])) for k in FIELDS[1:]}
        K.convec(dm, *(one[k] for k in FIELDS[1:]), conv, 1800.0, lmax)
        for k in FIELDS[1:]:
            assert_bitwise(D.download(one[k])[..., 0], D.download(dst[k])[..., t], f"{k} tracer {t}")
    sl = (slice(52, 71), slice(30, 42))  # the block the oracle runs (the operator is column-local)
    part = {k: np.ascontiguousarray(st[k][sl]) for k in FIELDS}
    c = {k: np.ascontiguousarray(getattr(met, k)[sl]) for k in ("entu", "detu", "entd", "detd", "cloud_top", "cloud_lfs")}
    convection.convec(*(part[k] for k in FIELDS), c["entu"], c["detu"], c["entd"], c["detd"], c["cloud_top"],
                      c["cloud_lfs"], np.ascontiguousarray(dkg[sl]), 1800.0, lmax)  # fmt: skip
    for k in FIELDS[1:]:
        assert_bitwise(D.download(dst[k])[sl], part[k], f"{k} against the oracle")


def test_advect_y_on_region1_halo(dev):
    """dynamv on the interior slice of region 1's halo'd bm (0:im+1, 0:jm+1, 0:lm+1): the
    faces outside it through the whole-layer scaling when a layer sub-cycles."""
    from oracle import advection
    from oracle.zoom import layer_halo

    D, K = dev
    m, b, q = records("y")[0]
    im, jm, lm = m.shape
    rng = np.random.default_rng(6)
    full = rng.uniform(-1.0, 1.0, (im + 2, jm + 2, lm + 2)) * np.abs(b).max()
    bm = full[1 : im + 1, 1 : jm + 2, 1 : lm + 1]  # data index = Fortran index (lower bound 0)
    bm[...] = 3.0 * b  # layers sub-cycle
    dm, dq, dfull = D.upload(m), [D.upload(x) for x in q], D.upload(full)
    dbm = dfull[1 : im + 1, 1 : jm + 2, 1 : lm + 1]
    nloop = K.advect_y(dm, *dq, dbm, full_bm=dfull, level_offset=1)
    layers = advection.advect_y(m, *q, bm, True)
    assert int(np.max(layers)) > 1  # some layers sub-cycle
    layer_halo(full, im, jm, layers)
    assert np.array_equal(nloop.cpu().numpy(), layers)
    for k, (g, w) in enumerate(zip([dm, *dq], [m, *q])):
        assert_bitwise(D.download(g), w, FIELDS[k])
    assert_bitwise(D.download(dfull), full, "full bm")


@pytest.mark.parametrize(("grid", "nt"), [
    ("small", 12),   # 10 levels
    ("glb3x2", 40),  # 34 levels: the apply's levels 32 and up, and two tracer ranges
])  # fmt: skip
def test_region_middle_split_form_equals_the_fused_step(dev, grid, nt):
    """Many tracers: an interval through the region driver (z fused into the one-kernel column
    form) and the same steps operator by operator ('V', the column middle without z, in the split
    form: the inverses through the context's buffer, then the tracers in 16-column workgroups,
    in tracer ranges), bit for bit."""
    import ctypes

    from pydyn_amd import _lib, meteo
    from pydyn_amd.region import Region

    D, K = dev
    print(f"split form against the fused step: {grid}, {nt} tracers", flush=True)
    met = meteo.Meteo(meteo.grid(grid), seed=3, nt=nt)
    fused, ops = Region(met), Region(met)
    ndyn, n = fused.interval()
    for s in range(n):
        for code, t in (("D", float(s * ndyn) + ndyn / 2.0), ("xyzVzyx", 0.0), ("M", 0.0)):
            _lib.check(_lib.lib().tm5_region_ops(ops._ptr, ctypes.c_double(t), int(ndyn), code.encode()), "tm5_region_ops")
    fused.check()
    ops.check()
    for k in ("m", "rm", "rxm", "rym", "rzm", "sp", "phlb"):
        assert_bitwise(D.download(getattr(ops, k)), D.download(getattr(fused, k)), k)
