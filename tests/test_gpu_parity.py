"""The HIP path against the reference's outputs, bit for bit.

Operator cases: the golden vectors the reference wrote (tests/golden/ops_tiny.npz), their
inputs uploaded, each operator run through libtm5hip, the outputs compared with the reference's.
Whole runs: the device region interval by interval against the reference's per-interval sha1s
(tests/golden/run_*_sha1.json) and, at glb3x2, against the oracle run on the same inputs.
"""

import hashlib
import json
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
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


def bits(x):
    return np.ascontiguousarray(x, dtype="<f8").view(np.int64)


def assert_bitwise(got, want, what=""):
    got, want = np.asarray(got, dtype=float), np.asarray(want, dtype=float)
    assert got.shape == want.shape, f"{what}: shape {got.shape} vs {want.shape}"
    differ = bits(got) != bits(want)
    if differ.any():
        first = tuple(int(i) for i in np.argwhere(differ)[0])
        raise AssertionError(f"{what}: {differ.sum()}/{differ.size} differ, first {first}: {got[first]!r} vs {want[first]!r}")


def case(name):
    ops = np.load(GOLDEN / "ops_tiny.npz")
    prefix = name + "."
    return {k[len(prefix) :]: ops[k] for k in ops.files if k.startswith(prefix)}


def sha(a):
    return hashlib.sha1(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()[:20]


# ------------------------------------------------------------------ operators
@pytest.mark.parametrize("axis", "xyz")
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_advection_matches_reference(dev, axis, seed):
    D, K = dev
    c = case(f"adv_{axis}_{seed}")
    print(f"advection {axis} case {seed}: shape {c['in_rm'].shape}", flush=True)
    st = {k: D.upload(c[f"in_{k}"]) for k in FIELDS}
    flux = D.upload(c["in_flux"])
    kern = {"x": K.advect_x, "y": K.advect_y, "z": K.advect_z}[axis]
    nloop = kern(*(st[k] for k in FIELDS), flux, bool(c["limits"]))
    for k in FIELDS:
        assert_bitwise(D.download(st[k]), c[f"out_{k}"], f"{axis} {k}")
    assert_bitwise(D.download(flux), c["out_flux"], f"{axis} flux")
    nloop = nloop.cpu().numpy() if hasattr(nloop, "cpu") else nloop
    assert int(np.max(nloop)) == int(c["nloop_max"])


def test_advection_cap_raises_cflerror(dev):
    D, K = dev
    c = case("adv_x_4")  # needs 19 sub-cycles
    st = {k: D.upload(c[f"in_{k}"]) for k in FIELDS}
    with pytest.raises(K.CFLError, match="dynamu: nloop too high"):
        K.advect_x(*(st[k] for k in FIELDS), D.upload(c["in_flux"]), True, max_nloop=5)
    c = case("adv_y_1")
    bm = c["in_flux"].copy()
    bm[:, 0] = 1.0
    st = {k: D.upload(c[f"in_{k}"]) for k in FIELDS}
    with pytest.raises(ValueError, match="polar"):
        K.advect_y(*(st[k] for k in FIELDS), D.upload(bm))


def test_dynam0_matches_reference(dev):
    D, K = dev
    c = case("dynam0")
    im, jm, lm = c["pv"].shape[0], c["pu"].shape[1], c["pu"].shape[2]
    am, bm, cm = D.empty((im + 1, jm, lm)), D.empty((im, jm + 1, lm)), D.empty((im, jm, lm + 1))
    K.fluxes_from_meteo(D.upload(c["pu"]), D.upload(c["pv"]), D.upload(c["bt"]), float(c["dt"]), am, bm, cm)
    assert_bitwise(D.download(am), c["am"], "am")
    assert_bitwise(D.download(bm), c["bm"], "bm")
    assert_bitwise(D.download(cm), c["cm"], "cm")


@pytest.mark.parametrize("name", ["conv_a", "conv_b", "conv_c", "conv_d"])
def test_convection_matches_reference(dev, name):
    D, K = dev
    import torch

    c = case(name)
    st = {k: D.upload(c[f"in_{k}"]) for k in FIELDS}
    conv = {k: D.upload(c[f"c_{k}"]) for k in ("entu", "detu", "entd", "detd", "dkg")}
    for k in ("cloud_top", "cloud_lfs"):
        conv[k] = D.upload(c[f"c_{k}"].astype(np.int32)).to(torch.int32)
    K.convec(*(st[k] for k in FIELDS), conv, 900.0, int(c["lmax"]))
    D.context().raise_errors()
    for k in FIELDS:
        assert_bitwise(D.download(st[k]), c[f"out_{k}"], f"{name} {k}")


def test_dkg_matches_reference(dev):
    D, K = dev
    c = case("dkg")
    out = K.dkg(D.upload(c["kvh"]), D.upload(c["m"]), D.upload(c["gph"]), int(c["lmax"]))
    assert_bitwise(D.download(out), c["dkg"], "dkg")


def test_decay_matches_reference(dev):
    D, K = dev
    c = case("decay")
    shape = (*c["loss"].shape[:3], 1)
    q = [D.upload(np.ones(shape)) for _ in range(4)]
    K.decay(D.upload(c["loss"]), float(c["dt"]), *q)
    for k in range(4):  # 1.0 * factor is the factor itself
        assert_bitwise(D.download(q[k])[..., 0], c["decay"], "decay")


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_mass_only_kernels_match_reference(dev, seed):
    D, K = dev
    c = case(f"cflm_{seed}")
    for axis, kern, fname, cap in (("x", K.mass_x_rows, "am", 10), ("y", K.mass_y, "bm", 1), ("z", K.mass_z, "cm", 1)):
        m, f = D.upload(c["m"]), D.upload(c[fname])
        ok = kern(m, f, max_nloop=cap) if axis != "z" else kern(m, f, iter_limit=cap)
        assert ok == bool(c[f"{axis}_ok"]), axis
        if ok:
            assert_bitwise(D.download(m), c[f"{axis}_m"], f"{axis} m")
            assert_bitwise(D.download(f), c[f"{axis}_flux"], f"{axis} flux")


def test_injection_matches_oracle_formula(dev):
    D, K = dev
    from oracle import physics

    rng = np.random.default_rng(3)
    im, jm, lm, nt = 8, 6, 5, 2
    rm, rzm = rng.uniform(0, 1, (im, jm, lm, nt)), rng.uniform(-1, 1, (im, jm, lm, nt))
    emis = rng.uniform(0, 1e3, (im, jm))
    factor = physics.month2dt(1800, 2592000.0)
    q = [D.upload(a) for a in (rm, np.zeros_like(rm), np.zeros_like(rm), rzm)]
    K.add_2d(D.upload(emis), factor, 1.0, 1.0, *q, level=2, tracer=2)
    x = physics.inject(emis, factor, 16.04275, 16.04275)
    physics.add_2d(rm, rzm, x, 2, 2)
    assert_bitwise(D.download(q[0]), rm, "rm")
    assert_bitwise(D.download(q[3]), rzm, "rzm")


# ------------------------------------------------------------------ whole runs
def _run_against_golden(name, limit=None):
    from pydyn_amd import meteo
    from pydyn_amd.region import Region

    ref = json.loads((GOLDEN / f"run_{name}_sha1.json").read_text())
    reg = Region(meteo.Meteo(meteo.grid(name), seed=ref["seed"], nt=ref["nt"]))
    for it, want in enumerate(ref["intervals"][:limit]):
        ndyn, n = reg.interval()
        assert (ndyn, n) == (want["ndyn"], want["n"]), f"{name} interval {it}"
        st = reg.state()
        for k, v in st.items():
            assert sha(v) == want[k], f"{name} interval {it}: {k} differs from the reference"
    return reg


def test_run_tiny_matches_reference(dev):
    reg = _run_against_golden("tiny")
    final = np.load(GOLDEN / "run_tiny_final.npz")
    st = reg.state()
    for k in final.files:
        assert_bitwise(st[k], final[k], k)


def test_run_small_with_cfl_cuts_matches_reference(dev):
    _run_against_golden("small")


def test_run_glb6x4_one_day_matches_reference(dev):
    _run_against_golden("glb6x4")


@pytest.mark.parametrize("name, batch", [("small", 40), ("small", 7), ("glb6x4", 24), ("glb3x2", 24)])
def test_queued_run_matches_reference(dev, name, batch):
    """Intervals queued in batches (tm5_region_run: the device read once per batch; each interval
    on the previous one's step, as checked steps that make Check_CFL's tests themselves; where a
    trial or a check disagrees, the interval's state restored from the backup and the run resumed
    from it with the reference's Check_CFL): the same (ndyn, n) and, at each batch's end, the
    same bits.  small turns from 3600 s to 1800 s (a checked step fails its test) and back (a
    trial expected to fail passes); glb3x2 x 34 does both 7 times or more in its 30 days."""
    from pydyn_amd import meteo
    from pydyn_amd.region import Region

    ref = json.loads((GOLDEN / f"run_{name}_sha1.json").read_text())
    reg = Region(meteo.Meteo(meteo.grid(name), seed=ref["seed"], nt=ref["nt"]))
    want = ref["intervals"]
    assert name != "small" or any(w["ndyn"] != 3600 for w in want)  # the batches meet Check_CFL cuts
    done = 0
    while done < len(want):
        count = min(batch, len(want) - done)
        got = reg.run(count)
        assert got == [(w["ndyn"], w["n"]) for w in want[done : done + count]], f"{name} intervals {done}.."
        done += count
        st = reg.state()
        for k, v in st.items():
            assert sha(v) == want[done - 1][k], f"{name} interval {done - 1}: {k} differs from the reference"


def test_run_glb3x2_matches_oracle_and_reference(dev):
    """glb3x2 x 34 (the headline grid): device and oracle bitwise equal for 3 intervals, and
    the reference's sha1s for as many intervals as the golden file holds (30 days when
    complete), checked daily."""
    from oracle.driver import Run
    from pydyn_amd import meteo
    from pydyn_amd.region import Region

    met = meteo.Meteo(meteo.grid("glb3x2"), seed=1)
    reg, run = Region(met), Run(met)
    for it in range(3):
        assert reg.interval() == run.interval(), f"interval {it}"
        st = reg.state()
        for k in st:
            assert_bitwise(st[k], run.st[k], f"glb3x2 interval {it}: {k}")
    path = GOLDEN / "run_glb3x2_sha1.json"
    if not path.exists():
        return
    ref = json.loads(path.read_text())["intervals"]
    reg = Region(met)
    for it, want in enumerate(ref):
        ndyn, n = reg.interval()
        assert (ndyn, n) == (want["ndyn"], want["n"]), f"glb3x2 interval {it}"
        if it % 24 == 23 or it == len(ref) - 1:
            st = reg.state()
            for k, v in st.items():
                assert sha(v) == want[k], f"glb3x2 interval {it}: {k} differs from the reference"
            print(f"glb3x2: interval {it + 1}/{len(ref)} bitwise equal to the reference", flush=True)


def test_tile_shapes_change_no_bits(dev):
    """Every tile shape of tm5_ctx_set_tile splits the same operations differently into
    workgroups: the small grid's run (Check_CFL cuts included) against the reference with shapes
    unlike the defaults and the tuned ones."""
    D, _ = dev
    ctx = D.context()
    saved = {k: ctx.tile(k) for k in D.TILES}
    try:
        for shape in ({"y_lines": 3, "z_lines": 5, "dynam0": 7, "m2phlb": 9},
                      {"y_lines": 16, "z_lines": 32, "dynam0": 64, "m2phlb": 64}):  # fmt: skip
            for k, v in shape.items():
                ctx.set_tile(k, v)
            _run_against_golden("small")
    finally:
        for k, v in saved.items():
            ctx.set_tile(k, v)
    with pytest.raises(RuntimeError):
        ctx.set_tile("dynam0", 65)


def test_profile_times_every_launch(dev):
    """profile_begin/profile_end total the GPU time per kernel from events around each launch."""
    D, _ = dev
    from pydyn_amd import meteo
    from pydyn_amd.region import Region

    reg = Region(meteo.Meteo(meteo.grid("tiny"), seed=3))
    ctx = D.context()
    ctx.profile_begin()
    reg.interval()
    prof = ctx.profile_end()
    # two x and two y kernels a step; a checked step searches nothing, its z sub-steps run inside
    # the column kernel, and its first x makes its fluxes (no Setup_MassFlow launch)
    for k, least in (("k_x<false, true>", 1), ("k_x<false, false>", 1), ("k_y_steps<false>", 2), ("k_m2phlb1", 1)):
        assert k in prof, (k, sorted(prof))
        count, secs = prof[k]
        assert count >= least and 0.0 < secs < 1.0, (k, prof[k])
    assert not {"k_y_search", "k_z_search", "k_z_steps<false>", "k_dynam0"} & set(prof), sorted(prof)
    assert any(k.startswith("k_column<") for k in prof), sorted(prof)
    reg.interval()  # not recorded
    assert ctx.profile_entries() == prof


@pytest.mark.parametrize(("scale", "nt"), [(8.0, 1), (9.0, 1), (9.0, 40)])
def test_queued_run_with_x_cuts_equals_the_reference_way(dev, scale, nt):
    """The zonal mean of the x mass flux scaled (a flow uniform along each row: no convergence, so
    y and z keep their Courant numbers) so that x rows need more than Check_CFL's 9 sub-cycles at
    3600 s in some intervals and not in others: the queued run's checked x tests fail inside the
    launch that runs a step's last x with the next step's first (either pass) and inside the
    batch's first x, and every failure is redone from the backup of its own interval.  The queued
    run equals the reference's order run interval by interval (tm5_region_interval: Check_CFL's
    rehearsal, the cut, then the steps) in every (ndyn, n) and every bit of the state; with 40
    tracers too (tracer ranges in the sweeps, the two x of a step pair on every tracer of a row)."""
    import ctypes

    from pydyn_amd import _lib, meteo
    from pydyn_amd.region import Region

    def met():
        m = meteo.Meteo(meteo.grid("small"), seed=5, nt=nt)
        zonal = m.pu0[:-1].mean(axis=0)  # (jm, lm)
        m.pu0 = m.pu0 + (scale - 1.0) * zonal[None, :, :]
        return m

    count = 120 if nt == 1 else 40  # the amplitude's 4-day period: 1 -> 2 steps, back, and (scale 9) again
    queued, plain = Region(met()), Region(met())
    got = queued.run(count)
    want = []
    nd, nn = ctypes.c_int(), ctypes.c_int()
    for k in range(count):
        rc = _lib.lib().tm5_region_interval(plain._ptr, 3600.0 * k, 3600, 3600, ctypes.byref(nd), ctypes.byref(nn))
        _lib.check(rc, "tm5_region_interval")
        want.append((nd.value, nn.value))
    assert got == want
    changes = sum(1 for a, b in zip(want, want[1:]) if a != b)
    assert changes >= (2 if nt == 1 else 1), want  # cut and uncut periods, entered (and left)
    a, b = queued.state(), plain.state()
    for k in a:
        assert sha(a[k]) == sha(b[k]), f"x scaled {scale}: {k} differs between the queued and the plain run"
