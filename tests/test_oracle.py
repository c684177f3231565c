"""The oracle against the reference's own outputs (golden vectors written by
tests/golden/make_golden.py) and against the reference tests' known answers and invariants.

Bitwise means equal float64 bit patterns (-0.0 differs from +0.0), the reference's definition
(pydyn/harness.py:150-153)."""

import hashlib
import json
import pathlib

import numpy as np
import pytest

from oracle import advection, cfl, convection, physics
from oracle import libm as oracle_libm
from oracle.driver import Run
from oracle.lines import nint
from pydyn_amd import meteo

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
OPS = np.load(GOLDEN / "ops_tiny.npz")


def bits(x):
    return np.ascontiguousarray(x, dtype="<f8").view(np.int64)


def assert_bitwise(got, want, what=""):
    got, want = np.asarray(got, dtype=float), np.asarray(want, dtype=float)
    assert got.shape == want.shape, what
    differ = bits(got) != bits(want)
    if differ.any():
        first = tuple(np.argwhere(differ)[0])
        raise AssertionError(f"{what}: {differ.sum()} of {differ.size} differ, first {first}: {got[first]!r} vs {want[first]!r}")


def case(name):
    prefix = name + "."
    return {k[len(prefix) :]: OPS[k] for k in OPS.files if k.startswith(prefix)}


FIELDS = ("m", "rm", "rxm", "rym", "rzm")


# ------------------------------------------------------------------ operators vs the reference
@pytest.mark.parametrize("axis", "xyz")
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_advection_bitwise(axis, seed):
    c = case(f"adv_{axis}_{seed}")
    st = {k: c[f"in_{k}"].copy() for k in FIELDS}
    flux = c["in_flux"].copy()
    kern = {"x": advection.advect_x, "y": advection.advect_y, "z": advection.advect_z}[axis]
    nloop = kern(*(st[k] for k in FIELDS), flux, bool(c["limits"]))
    for k in FIELDS:
        assert_bitwise(st[k], c[f"out_{k}"], f"{axis} {k}")
    assert_bitwise(flux, c["out_flux"], f"{axis} flux")
    assert int(np.max(nloop)) == int(c["nloop_max"])


def test_dynam0_bitwise():
    c = case("dynam0")
    am, bm, cm = advection.dynam0(c["pu"], c["pv"], c["bt"], float(c["dt"]))
    assert_bitwise(am, c["am"], "am")
    assert_bitwise(bm, c["bm"], "bm")
    assert_bitwise(cm, c["cm"], "cm")


@pytest.mark.parametrize("name", ["conv_a", "conv_b", "conv_c", "conv_d"])
def test_convection_bitwise(name):
    c = case(name)
    lmax = int(c["lmax"])
    st = {k: c[f"in_{k}"].copy() for k in FIELDS}
    meteo_ = {k: c[f"c_{k}"] for k in ("entu", "detu", "entd", "detd", "cloud_top", "cloud_lfs", "dkg")}
    ncol = st["m"].shape[0] * st["m"].shape[1]
    conv1 = convection.convection_matrix(
        st["m"].reshape(ncol, -1)[:, :lmax], *(meteo_[k].reshape(ncol, -1) for k in ("entu", "detu", "entd", "detd")),
        meteo_["cloud_top"].ravel(), meteo_["cloud_lfs"].ravel(), meteo_["dkg"].reshape(ncol, -1), 900.0, lmax,
    )  # fmt: skip
    assert_bitwise(conv1, c["conv1"], "conv1")
    assert_bitwise(convection.crout_inverse(conv1), c["inverse"], "inverse")
    convection.convec(*(st[k] for k in FIELDS), *meteo_.values(), 900.0, lmax)
    for k in FIELDS:
        assert_bitwise(st[k], c[f"out_{k}"], k)


def test_dkg_bitwise():
    c = case("dkg")
    assert_bitwise(physics.dkg(c["kvh"], c["m"], c["gph"], int(c["lmax"])), c["dkg"], "dkg")


def test_decay_bitwise():
    c = case("decay")
    assert_bitwise(physics.decay(c["loss"], float(c["dt"])), c["decay"], "decay")


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_mass_only_kernels(seed):
    c = case(f"cflm_{seed}")
    for axis, kern, fname, cap in (("x", cfl.mass_x, "am", 10), ("y", cfl.mass_y, "bm", 1), ("z", cfl.mass_z, "cm", 1)):
        m, f = c["m"].copy(), c[fname].copy()
        ok = kern(m, f, cap)
        assert ok == bool(c[f"{axis}_ok"]), axis
        if ok:
            assert_bitwise(m, c[f"{axis}_m"], f"{axis} m")
            assert_bitwise(f, c[f"{axis}_flux"], f"{axis} flux")


# ------------------------------------------------------------------ whole runs vs the reference
def sha(a):
    return hashlib.sha1(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()[:20]


def run_matches(name, limit=None):
    ref = json.loads((GOLDEN / f"run_{name}_sha1.json").read_text())
    run = Run(meteo.Meteo(meteo.grid(name), seed=ref["seed"], nt=ref["nt"]))
    for it, want in enumerate(ref["intervals"][:limit]):
        ndyn, n = run.interval()
        assert (ndyn, n) == (want["ndyn"], want["n"]), f"interval {it}"
        for k, v in run.st.items():
            if k in want:
                assert sha(v) == want[k], f"{name} interval {it}: {k}"
    return run


def test_run_tiny_bitwise():
    run = run_matches("tiny")
    final = np.load(GOLDEN / "run_tiny_final.npz")
    for k in final.files:
        assert_bitwise(run.st[k], final[k], k)


def test_run_small_with_cfl_cuts_bitwise():
    ref = json.loads((GOLDEN / "run_small_sha1.json").read_text())
    assert any(i["ndyn"] != 3600 for i in ref["intervals"])  # the fixture covers Check_CFL cutting ndyn
    run_matches("small")


def test_run_glb6x4_one_day_bitwise():
    run_matches("glb6x4")


def test_run_glb3x2_first_intervals_bitwise():
    path = GOLDEN / "run_glb3x2_sha1.json"
    if not path.exists():
        pytest.skip("no glb3x2 golden run")
    run_matches("glb3x2", limit=2)


# ------------------------------------------------------------------ known answers (reference tests)
def test_new_valid_timestep_known_answers():
    # tests/test_cfl.py:25-30
    assert cfl.new_valid_timestep(3600, 3600, 3600) == 1800
    assert cfl.new_valid_timestep(1800, 3600, 3600) == 1200
    assert cfl.new_valid_timestep(900, 3600, 3600, tref_max=2) == 720
    with pytest.raises(ValueError):
        cfl.new_valid_timestep(1, 3600, 3600)


def test_nint_is_half_away_from_zero():
    # tests/test_cfl.py:84-92: n = 3 at 1800 s reduced to 1200 s is 4.5 steps -> 5
    assert nint(float(3 * 1800) / float(1200)) == 5
    assert nint(2.5) == 3 and nint(-2.5) == -3 and nint(0.49) == 0


def test_month2dt_and_xmch4_known_answers():
    # tests/test_physics.py:149-152
    assert physics.month2dt(3600, 2678400.0) == 1800.0 / 2678400.0
    assert physics.month2dt(3600, 2678400.0, tref=2) == 900.0 / 2678400.0
    assert meteo.XMCH4 == 16.04275


def test_check_cfl_cuts_and_rounds():
    """A flow that fails the y cap at 3600 s passes at 1800 s: (1800, 2) and am/bm/cm scaled."""
    met = meteo.Meteo(meteo.grid("small"), seed=5)
    run = Run(met)
    t0 = 2 * 86400.0  # the amplitude's peak
    am0 = run.st["am"].copy()
    ndyn, n, it = cfl.check_cfl(run.st, lambda i, nd: met.fluxes(t0 + (i - 0.5) * nd), 3600, 1, 3600, met.g.bt)
    assert (ndyn, n, it) == (1800, 2, 2)
    assert_bitwise(run.st["am"], am0 * 0.5, "am scaled by ndyn_new / ndyn_old")


# ------------------------------------------------------------------ invariants
def test_crout_inverse_is_an_inverse():
    # tests/test_kernels.py:174-189: fastminv ~ numpy.linalg.inv
    c = case("conv_a")
    np.testing.assert_allclose(convection.crout_inverse(c["conv1"]), np.linalg.inv(c["conv1"]), rtol=1e-10, atol=1e-14)


@pytest.mark.parametrize("axis", "xyz")
def test_advection_conserves_mass_and_uniform_ratio(axis):
    # tests/test_kernels.py:23-46
    c = case(f"adv_{axis}_1")
    m = c["in_m"].copy()
    rm = np.stack([m * 1e-6, m * 3e-6], axis=-1)
    z = np.zeros_like(rm)
    st = [m, rm, z.copy(), z.copy(), z.copy()]
    m0, r0 = m.sum(), rm.sum(axis=(0, 1, 2))
    kern = {"x": advection.advect_x, "y": advection.advect_y, "z": advection.advect_z}[axis]
    kern(*st, c["in_flux"].copy())
    np.testing.assert_allclose(st[0].sum(), m0, rtol=1e-13)
    np.testing.assert_allclose(st[1].sum(axis=(0, 1, 2)), r0, rtol=1e-13)
    np.testing.assert_allclose(st[1] / st[0][..., None], [1e-6, 3e-6] * np.ones_like(st[1]), rtol=1e-12)


def test_fdlibm_exp_within_one_ulp_of_the_c_library():
    import math

    x = np.concatenate([np.linspace(-700.0, 700.0, 20001), np.linspace(-0.4, 0.4, 20001)])
    got = oracle_libm.exp(x)
    want = np.array([math.exp(v) for v in x])
    assert np.all(np.abs(got - want) <= np.spacing(want))
    special = oracle_libm.exp(np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 710.0, -746.0]))
    assert special[0] == 1.0 and special[1] == 1.0 and special[2] == np.inf and special[3] == 0.0
    assert np.isnan(special[4]) and special[5] == np.inf and special[6] == 0.0


def _ulps(got, want):
    """|got - want| in units in the last place of want (finite, same sign)."""
    return np.abs(got.view(np.int64) - want.view(np.int64))


def test_fdlibm_log_within_one_ulp_of_the_c_library():
    import math

    rng = np.random.default_rng(11)
    x = np.concatenate([
        np.exp(rng.uniform(-744.0, 709.0, 100000)),  # every binade, subnormals included
        rng.uniform(0.5, 2.0, 100000),  # around 1, both of fdlibm's branches
        1.0 + rng.uniform(-1e-6, 1e-6, 20000),  # |f| < 2^-20
        np.ldexp(1.0, np.arange(-1074, 1024)),  # powers of two: k ln2 exactly
    ])  # fmt: skip
    got = oracle_libm.log(x)
    want = np.array([math.log(v) for v in x])
    assert _ulps(got, want).max() <= 1
    assert oracle_libm.log(np.array([2.0]))[0] == math.log(2.0)  # ln2_hi + ln2_lo rounds to ln 2
    special = oracle_libm.log(np.array([1.0, 0.0, -0.0, np.inf, -1.0, np.nan, -np.inf]))
    assert special[0] == 0.0 and special[1] == -np.inf and special[2] == -np.inf and special[3] == np.inf
    assert np.isnan(special[4:]).all()


def test_fdlibm_pow_within_one_ulp_of_the_c_library_and_c99_special_cases():
    rng = np.random.default_rng(12)
    x = np.concatenate([np.exp(rng.uniform(-30.0, 30.0, 100000)), rng.uniform(0.1, 3.0, 100000),
                        1.0 + rng.uniform(-1e-7, 1e-7, 2000), -np.exp(rng.uniform(-5.0, 5.0, 2000))])  # fmt: skip
    y = np.concatenate([rng.uniform(-20.0, 20.0, 100000), rng.uniform(-3.0, 3.0, 100000),
                        rng.uniform(-3e9, 3e9, 2000), np.round(rng.uniform(-20.0, 20.0, 2000))])  # fmt: skip
    got = oracle_libm.pow(x, y)
    with np.errstate(all="ignore"):
        want = np.power(x, y)
    fine = np.isfinite(want) & (want != 0.0)
    # |y| > 2^31 with x within 2^-20 of 1 takes fdlibm's short series for log2(x), whose head
    # ivln2_h * t is not exact (21 x 33 bits): up to a few hundred ulps there, by the method
    corner = (np.abs(y) > 2.0**31) & (np.abs(x - 1.0) < 2.0**-20)
    assert _ulps(got[fine & ~corner], want[fine & ~corner]).max() <= 1
    assert np.all(np.abs(got[fine & corner] / want[fine & corner] - 1.0) < 1e-12)
    assert np.array_equal(got[~fine], want[~fine])
    k = np.arange(-1074, 1024, dtype=np.float64)
    assert np.array_equal(oracle_libm.pow(2.0, k), np.ldexp(1.0, k.astype(int)))  # exact powers of two
    v = rng.uniform(0.0, 10.0, 1000)
    assert np.array_equal(oracle_libm.pow(v, 2.0), v * v)
    assert np.array_equal(oracle_libm.pow(v, 0.5), np.sqrt(v))
    assert np.array_equal(oracle_libm.pow(v, -1.0), 1.0 / v)
    inf, nan = np.inf, np.nan
    cases = [  # C99 Annex F.9.4.4
        (nan, 0.0, 1.0), (1.0, nan, 1.0), (-1.0, inf, 1.0), (-1.0, -inf, 1.0), (-2.0, 3.0, -8.0),
        (0.0, -1.0, inf), (-0.0, -1.0, -inf), (-0.0, -2.0, inf), (0.0, 3.0, 0.0), (-0.0, 3.0, -0.0),
        (0.5, inf, 0.0), (0.5, -inf, inf), (2.0, -inf, 0.0), (2.0, inf, inf), (-inf, -3.0, -0.0),
        (-inf, 3.0, -inf), (-inf, 2.0, inf), (inf, -1.0, 0.0), (2.0, 1024.0, inf), (2.0, -1075.0, 0.0),
        (2.0, -1074.0, 5e-324), (-8.0, 1.0 / 3.0, nan), (nan, 1.0, nan), (2.0, nan, nan),
    ]  # fmt: skip
    for a, b, want in cases:
        g = oracle_libm.pow(a, b)[()]
        assert (np.isnan(g) and np.isnan(want)) or (g == want and np.signbit(g) == np.signbit(want)), (a, b, g)


# ------------------------------------------------------------------ the zoom tree
def _zoom_against_golden(name, limit=None):
    import hashlib
    import json

    from oracle import zoom as OZ
    from pydyn_amd import meteo

    ref = json.loads((GOLDEN / f"zoom_{name}_sha1.json").read_text())["intervals"][:limit]
    run = OZ.ZoomRun(*meteo.zoom(name), redgrid=meteo.redgrid(name))
    for it, want in enumerate(ref):
        assert run.interval() == (want["ndyn"], want["n"]), f"{name} interval {it}"
        for key, digest in want.items():
            if "." in key:
                r, k = int(key[1]), key.split(".")[1]
                got = hashlib.sha1(np.ascontiguousarray(run.st[r][k].data, dtype="<f8").tobytes()).hexdigest()[:20]
                assert got == digest, f"{name} interval {it}: region {r} {k}"
    return run


def test_zoom_tiny_tree_bitwise():
    """A global region with a nested child, 4 simulated days: every array of both regions (halos
    included) bitwise equal to the reference's zoom path, interval by interval."""
    run = _zoom_against_golden("tiny2")
    final = np.load(GOLDEN / "zoom_tiny2_final.npz")
    for key in final.files:
        r, k = int(key[1]), key.split(".")[1]
        assert np.array_equal(run.st[r][k].data.view(np.int64), final[key].view(np.int64)), key


def test_zoom_small_tree_with_cfl_cuts_bitwise():
    run = _zoom_against_golden("small2", limit=36)
    assert any(nd == 1800 for _, nd, _ in run.log), "the sample must include a Check_CFL cut"


def test_zoom_production_tree_first_intervals_bitwise():
    """glb3x2 with a 1x1 degree child over 170W-50W, 10N-80N (120 x 70 x 34)."""
    _zoom_against_golden("c2", limit=2)


def test_zoom_reduced_grid_tiny_tree_bitwise():
    """Region 1's x advection on a reduced grid (pydyn/redgrid.py), 4 simulated days."""
    _zoom_against_golden("tiny2r")


def test_redgrid_bands_known_answer():
    """The reference test's production rc (tests/test_redgrid.py:25-37): glb6x4 rows 1, 2, 3, 45
    merged 60, 20, 10, 60."""
    from oracle.zoom import RedGrid

    g = RedGrid(60, 45, (1, 2, 3, 45), (60, 20, 10, 60))
    assert g.bands == [(0, 60, 1), (1, 20, 3), (2, 10, 6), (44, 60, 1)]
    assert g.imred[3] == 60 and g.imred[44] == 1


# ------------------------------------------------------------------ the tangent-linear and adjoint models
def _window(name, intervals, limits, clipped=True, seed=3, **kw):
    from oracle import adjoint as A

    met = meteo.Meteo(meteo.grid(name), seed=seed)
    rm = met.state()["rm"]
    rng = np.random.default_rng(5)
    init = [rm] + [(2.0 * rm * rng.uniform(-1.0, 1.0, rm.shape) if clipped else np.zeros_like(rm)) for _ in range(3)]
    return met, init, A.Window(met, intervals, limits=limits, init=init, **kw)


def _directions(met, shape, seed):
    rng = np.random.default_rng(seed)
    rm = met.state()["rm"]
    return [rm * rng.standard_normal(shape) for _ in range(4)], met.emis * rng.standard_normal(met.emis.shape)


def test_tangent_is_the_forward_when_linear():
    """Limiter off, the tracer map is linear in (initial tracers, emission): the forward run from
    the perturbation itself (the forward the golden vectors pin) equals the tangent."""
    from oracle import adjoint as A

    met, init, W = _window("small", 3, limits=False)
    dx, de = _directions(met, W.final[0].shape, 21)
    tl = W.tlm(dx, de)
    F = A.Window(met, 3, limits=False, init=dx, emis=de).final
    for k in range(4):
        assert np.abs(tl[k] - F[k]).max() <= 1e-13 * np.abs(F[k]).max(), A.FIELDS[k]


@pytest.mark.parametrize("limits", [True, False])
@pytest.mark.parametrize("name", ["tiny", "small"])
def test_adjoint_is_the_transpose_of_the_tangent(name, limits):
    met, init, W = _window(name, 3, limits)
    dx, de = _directions(met, W.final[0].shape, 22)
    lam = [np.random.default_rng(23).standard_normal(W.final[0].shape) for _ in range(4)]
    tl = W.tlm(dx, de)
    ad, ae = W.adm(lam)
    lhs = sum(float(np.sum(a * b)) for a, b in zip(tl, lam))
    rhs = sum(float(np.sum(a * b)) for a, b in zip(dx, ad)) + float(np.sum(de * ae))
    assert abs(lhs - rhs) <= 1e-13 * abs(lhs)


def test_tangent_matches_finite_differences_with_the_limiter_clipping():
    """Central differences of the forward with slopes the limiter clips: the tangent freezes the
    limiter's branches, which small perturbations keep."""
    from oracle import adjoint as A

    met, init, W = _window("small", 3, limits=True)
    dx, de = _directions(met, W.final[0].shape, 24)
    clipped = sum(int((A.lim_branch(Q[1], Q[0]) != 0).sum()) for op, _, _, _, Q in W.ops if op in "xyz")
    assert clipped > 1000
    tl = W.tlm(dx, de)
    eps = 1e-7
    plus = A.Window(met, 3, init=[x + eps * d for x, d in zip(init, dx)], emis=met.emis + eps * de).final
    minus = A.Window(met, 3, init=[x - eps * d for x, d in zip(init, dx)], emis=met.emis - eps * de).final
    for k in range(4):
        fd = (plus[k] - minus[k]) / (2.0 * eps)
        assert np.abs(fd - tl[k]).max() <= 1e-6 * np.abs(tl[k]).max(), A.FIELDS[k]


@pytest.mark.parametrize("axis", "xyz")
def test_operator_adjoints_with_sub_cycling(axis):
    """Each advection operator's adjoint is the transpose of its tangent when its fluxes are
    scaled up so that it sub-cycles (x 3, y 3, z 8 times the window's fluxes)."""
    from oracle import adjoint as A
    from oracle.advection import closed_search, periodic_search

    met, init, W = _window("small", 1, limits=True)
    op_fn = {"x": A.x_op, "y": A.y_op, "z": A.z_op}[axis]
    scale = {"x": 3.0, "y": 3.0, "z": 8.0}[axis]
    rng = np.random.default_rng(25)
    first_seen = False
    for op, _, m, flux, Q in W.ops:
        if op != axis:
            continue
        f = flux * scale
        if axis == "x":
            n = int(periodic_search(m.reshape(m.shape[0], -1).copy(), f.reshape(m.shape[0] + 1, -1).copy(), 49).max())
        elif axis == "y":
            n = max(closed_search(m[:, :, k].T.copy(), f[:, :, k].T.copy(), 9, fresh=False)[0] for k in range(m.shape[2]))
        else:
            n = closed_search(np.moveaxis(m, 2, 0).copy(), np.moveaxis(f, 2, 0).copy(), 15, fresh=True)[0]
        assert n >= 2
        d = [rng.standard_normal(Q[0].shape) for _ in range(4)]
        lam = [rng.standard_normal(Q[0].shape) for _ in range(4)]
        tl = op_fn(m, f, Q, d, True, "tl")
        ad = op_fn(m, f, Q, lam, True, "ad")
        lhs = sum(float(np.sum(a * b)) for a, b in zip(tl, lam))
        rhs = sum(float(np.sum(a * b)) for a, b in zip(d, ad))
        assert abs(lhs - rhs) <= 1e-13 * abs(lhs)
        # and the tangent against central differences of the forward operator, on the step's
        # first operator of the axis: later ones can be handed slopes exactly at the limiter's
        # bound (a clip carried through physics that scales s and rm alike), a kink where the
        # one-sided derivatives differ and the tangent takes the forward's branch
        if first_seen:
            continue
        first_seen = True
        forward = {"x": advection.advect_x, "y": advection.advect_y, "z": advection.advect_z}[axis]
        eps = 1e-6 * float(np.abs(Q[0]).max())
        out = []
        for sign in (1.0, -1.0):
            st = [m.copy(), *(q + sign * eps * x for q, x in zip(Q, d)), f.copy()]
            forward(*st)
            out.append(st[1:5])
        for k in range(4):
            fd = (out[0][k] - out[1][k]) / (2.0 * eps)
            assert np.abs(fd - tl[k]).max() <= 1e-6 * np.abs(tl[k]).max(), (axis, k)


# ------------------------------------------------------------------ CH4 loss rates
def _chem_case(case):
    g = np.load(GOLDEN / "chem_loss.npz")
    pre = f"{case}."
    return {k[len(pre):]: g[k] for k in g.files if k.startswith(pre)}


def _loss_inputs(c):
    keys = ("temper", "phlb", "lat", "strat_press", "strat_oh", "strat_cl", "strat_od", "oh_3d", "pl_trop",
            "cl_trop", "soil")  # fmt: skip
    kw = {k: c[k] for k in keys if k in c}
    for k in ("oh_scale", "cl_scale", "o1d_scale", "soil_rate"):
        if k in c:
            kw[k] = float(c[k])
    return kw


@pytest.mark.parametrize("case", ["plain", "trop", "high"])
def test_loss_rates_and_decay_match_the_reference(case):
    from oracle import chem

    c = _chem_case(case)
    loss = chem.loss_rates(**_loss_inputs(c))
    assert np.array_equal(loss.view(np.int64), c["loss"].view(np.int64))
    assert np.array_equal(chem.decay(loss, 1800.0).view(np.int64), c["decay"].view(np.int64))


def test_loss_rate_known_answers():
    """The reference tests' known answers (tests/test_physics.py:97-101): edges of the zonal
    columns, the tropopause at the equator and the pole; and a rebin whose means are exact."""
    from oracle import chem

    assert np.array_equal(chem.strat_edges(np.array([[900.0, 700.0, 300.0]])), [[100000.0, 80000.0, 50000.0, 0.0]])
    assert chem.tropopause(np.array([0.0]))[0] == 30000.0 - 21500.0
    assert abs(chem.tropopause(np.array([np.pi / 2]))[0] - 30000.0) < 1e-9
    # layers [1000, 600] and [600, 0] of values 2 and 5 onto [1000, 800, 400, 0]: means 2, (2*200 + 5*200)/400, 5
    got = chem.rebin(np.array([1000.0, 600.0, 0.0]), np.array([2.0, 5.0]), np.array([1000.0, 800.0, 400.0, 0.0]))
    assert np.array_equal(got, [2.0, 3.5, 5.0])
    with pytest.raises(ValueError):
        chem.rebin(np.array([0.0, 600.0, 1000.0]), np.array([2.0, 5.0]), np.array([1000.0, 0.0]))


# ------------------------------------------------------------------ Calc_Kzz
def kzz_case(case):
    g = np.load(GOLDEN / "kzz.npz")
    c = {k[len(case) + 1:]: g[k] for k in g.files if k.startswith(case + ".")}
    surface = {k[len("surface_"):]: c[k] for k in c if k.startswith("surface_")}
    args = (c["spm"], c["sp"], c["pu"], c["pv"], c["t"], c["q"], surface, c["at"], c["bt"], c["dxyp"],
            int(c["lmax_conv"]))  # fmt: skip
    grid = {"dx": float(c["dx"]), "dy": float(c["dy"]), "ybeg": float(c["ybeg"])}
    return args, grid, c


@pytest.mark.parametrize("case", ["a", "b"])
def test_calc_kzz_matches_the_reference(case):
    from oracle import diffusion as OD

    args, grid, c = kzz_case(case)
    dkg, parts = OD.calc_kzz(*args, **grid)
    for k, v in {"dkg": dkg, **parts}.items():
        assert np.array_equal(v.view(np.int64), c[k].view(np.int64)), k


def test_calc_kzz_known_answers():
    """The reference tests' known answers (tests/test_diffusion.py:20-69): jqif on the sign bit,
    ustar blending land and sea by the land fraction, hydrostatic interface heights, dkg closed
    at lmax_conv and non-negative."""
    from oracle import diffusion as OD

    x = np.array([1.0, 0.0, -0.0, -1.0, 1e-300])
    assert OD.jqif(x, 0.0).tolist() == [1.0, 1.0, 0.0, 0.0, 1.0]
    assert OD.jqif(0.0, x).tolist() == [0.0, 1.0, 1.0, 1.0, 0.0]
    rng = np.random.default_rng(16)
    lsm = np.array([[0.0, 100.0, 40.0]])
    f = [rng.uniform(0.5, 1.0, (1, 3)) for _ in range(4)]
    fields = (0.01 * f[0], 0.5 * f[1], 3.0 * f[2], -2.0 * f[3], 0.1 * f[0], 0.2 * f[1])
    u, sea, land = OD.ustar(lsm, *fields), OD.ustar(np.zeros((1, 3)), *fields), OD.ustar(np.full((1, 3), 100.0), *fields)
    assert u[0, 0] == sea[0, 0] and u[0, 1] == land[0, 1]
    assert u[0, 2] == 0.4 * land[0, 2] + (1.0 - 0.4) * sea[0, 2]
    sp = rng.uniform(9.0e4, 1.05e5, (4, 3))
    a, b = 1.0e4 * np.sin(np.linspace(0.0, np.pi, 9)), np.linspace(1.0, 0.0, 9)
    t, q = np.full((4, 3, 8), 250.0), np.zeros((4, 3, 8))
    oro = rng.uniform(0.0, 1.0e4, (4, 3))
    gph = OD.geopotential(oro, sp, t, q, a, b)
    phlb = a + b * sp[..., None]
    steps = 250.0 * 287.05 * np.log(phlb[..., :-2] / phlb[..., 1:-1]) / OD.GRAV
    np.testing.assert_allclose(np.diff(gph[..., :-1], axis=-1), steps, rtol=1e-13)
    assert np.array_equal(gph[..., 0], oro / OD.GRAV) and (gph[..., -1] == 200000.0).all()
    args, grid, c = kzz_case("b")
    dkg, _ = OD.calc_kzz(*args, **grid)
    lmax = dkg.shape[-1]
    assert np.isfinite(dkg).all() and (dkg >= 0.0).all() and (dkg[..., lmax - 1] == 0.0).all()
