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
