"""The device adjoint against the NumPy reference adjoint (oracle/adjoint.py), and by the
dot-product and finite-difference tests of the forward model.

Operator cases take the forward inputs an oracle window recorded (with initial slopes large
enough that the limiter clips, and the fluxes scaled up so that every axis sub-cycles) and
compare the device's lambdas with the oracle's.  Window cases compare whole adjoints, and at
the 120 x 90 x 34 grid over 24 hourly intervals run the dot-product test (limiter off: the map
is then linear, so the forward run is its own tangent) and finite differences of the loss
through the torch.autograd.Function (limiter on).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("rm", "rxm", "rym", "rzm")
SCALE = {"x": 3.0, "y": 3.0, "z": 8.0}  # flux factors under which every axis sub-cycles


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from pydyn_amd import adjoint as AD
    from pydyn_amd import device as D

    D.context().flags(clear=True)
    return D, AD


def rel(got, want):
    scale = float(np.abs(want).max())
    return float(np.abs(np.asarray(got) - want).max()) / (scale if scale > 0 else 1.0)


def clipped_init(met, seed=5):
    """The meteo's rm with slopes of up to twice |rm|: the limiter clips many of them."""
    rng = np.random.default_rng(seed)
    rm = met.state()["rm"]
    return [rm] + [2.0 * rm * rng.uniform(-1.0, 1.0, rm.shape) for _ in range(3)]


def oracle_window(name, intervals, limits=True, seed=3):
    from oracle import adjoint as A
    from pydyn_amd import meteo

    met = meteo.Meteo(meteo.grid(name), seed=seed)
    return met, A.Window(met, intervals, limits=limits, init=clipped_init(met))


# ------------------------------------------------------------------ operators
@pytest.mark.parametrize("limits", [True, False])
@pytest.mark.parametrize("scaled", [False, True])
@pytest.mark.parametrize("axis", "xyz")
def test_operator_adjoint_matches_oracle(dev, axis, scaled, limits):
    from oracle import adjoint as A

    D, AD = dev
    scale = SCALE[axis] if scaled else 1.0
    print(f"operator adjoint {axis} scale {scale} limits {limits}", flush=True)
    met, W = oracle_window("small", 1)
    rng = np.random.default_rng(11)
    oracle_op = {"x": A.x_op, "y": A.y_op, "z": A.z_op}[axis]
    cases = [rec for rec in W.ops if rec[0] == axis]
    assert len(cases) == 2
    for op, ndyn, m, flux, Q in cases:
        flux = flux * scale
        lam = [rng.standard_normal(Q[0].shape) for _ in range(4)]
        want = oracle_op(m, flux, Q, lam, limits, "ad")
        q_dev = [D.upload(x) for x in Q]
        lam_dev = [D.upload(x) for x in lam]
        AD.advect_adjoint(axis, D.upload(m), D.upload(flux), q_dev, lam_dev, limits)
        D.context().raise_errors()
        for k, (g, w) in enumerate(zip(lam_dev, want)):
            assert rel(D.download(g), w) <= 1e-12, f"{axis} field {FIELDS[k]}: {rel(D.download(g), w):.2e}"
        for x, y in zip(q_dev, Q):  # the forward inputs are left alone
            assert np.array_equal(D.download(x), y)


def test_column_adjoint_matches_oracle(dev):
    D, AD = dev
    print("column adjoint small", flush=True)
    met, W = oracle_window("small", 1)
    window = AD.Window(met, 1)
    rng = np.random.default_rng(12)
    rec = next(r for r in W.ops if r[0] == "v")
    _, ndyn, m, _, _ = rec
    lam = [rng.standard_normal(W.final[0].shape) for _ in range(4)]
    lam_emis = np.zeros_like(met.emis)
    want = [x.copy() for x in lam]
    for op in "vscMcsv"[::-1]:  # the transposes in reverse
        if op != "M":
            want = W._op(op, ndyn, m, None, None, want, "ad", lam_emis)
    lam_dev = [D.upload(x) for x in lam]
    le = D.zeros(met.emis.shape)
    window.column_adjoint(D.upload(m), ndyn, lam_dev, le)
    D.context().raise_errors()
    for k, (g, w) in enumerate(zip(lam_dev, want)):
        assert rel(D.download(g), w) <= 1e-12, f"{FIELDS[k]}"
    assert rel(D.download(le), lam_emis) <= 1e-12
    assert np.abs(lam_emis).max() > 0.0


# ------------------------------------------------------------------ windows
@pytest.mark.parametrize(("grid", "intervals", "limits"), [
    ("small", 3, True),
    ("small", 3, False),
    ("glb6x4", 24, True),  # 60 x 45 x 25 over a day, slopes the limiter clips
])  # fmt: skip
def test_window_adjoint_matches_oracle(dev, grid, intervals, limits):
    """The device adjoint of a window against the NumPy adjoint of the same (bit-identical)
    forward, to 1e-11 relative."""
    D, AD = dev
    print(f"window adjoint {grid} x {intervals} limits {limits}", flush=True)
    met, W = oracle_window(grid, intervals, limits=limits)
    window = AD.Window(met, intervals, limits=limits)
    init = [D.upload(x) for x in clipped_init(met)]
    final = window.forward(init, D.upload(met.emis))
    for k, (g, w) in enumerate(zip(final, W.final)):
        assert np.array_equal(D.download(g), w), f"forward {FIELDS[k]}"
    rng = np.random.default_rng(13)
    lam = [rng.standard_normal(W.final[0].shape) for _ in range(4)]
    want, want_e = W.adm(lam)
    got, got_e = window.backward([D.upload(x) for x in lam])
    for k, (g, w) in enumerate(zip(got, want)):
        assert rel(D.download(g), w) <= 1e-11, f"{FIELDS[k]}: {rel(D.download(g), w):.2e}"
    assert rel(D.download(got_e), want_e) <= 1e-11


def test_replay_reproduces_the_forward(dev):
    """Each interval replayed from its checkpoint ends on the next checkpoint, bit for bit."""
    D, AD = dev
    from pydyn_amd import meteo

    met = meteo.Meteo(meteo.grid("small"), seed=3)
    window = AD.Window(met, 3)
    final = window.forward([D.upload(x) for x in clipped_init(met)], D.upload(met.emis))
    for k in range(3):
        window._replay(k)
        after = window.checkpoints[k + 1][3] if k < 2 else dict(zip(FIELDS, final), m=window.m_final)
        for name, want in after.items():
            assert D.download(getattr(window.region, name)).tobytes() == D.download(want).tobytes(), (k, name)


def test_mixed_dot_product_numpy_tangent_device_adjoint(dev):
    """<TLM d, l> from the NumPy tangent-linear model = <d, ADM l> from the device, limiter on."""
    D, AD = dev
    print("mixed dot product glb6x4", flush=True)
    met, W = oracle_window("glb6x4", 3)
    rng = np.random.default_rng(14)
    shape = W.final[0].shape
    scale = float(np.abs(W.ops[0][4][0]).mean())
    dx = [1e-3 * scale * rng.standard_normal(shape) for _ in range(4)]
    de = 1e-3 * float(np.abs(met.emis).mean()) * rng.standard_normal(met.emis.shape)
    lam = [rng.standard_normal(shape) for _ in range(4)]
    tl = W.tlm(dx, de)
    window = AD.Window(met, 3)
    window.forward([D.upload(x) for x in clipped_init(met)], D.upload(met.emis))
    ad, ae = window.backward([D.upload(x) for x in lam])
    lhs = sum(float(np.sum(a * b)) for a, b in zip(tl, lam))
    rhs = sum(float(np.sum(a * D.download(b))) for a, b in zip(dx, ad)) + float(np.sum(de * D.download(ae)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def test_replay_after_a_second_forward_uses_its_own_trajectory(dev):
    """Two forwards of one window before the first graph's backward: that backward replays the
    first forward's trajectory (its own checkpoints), giving the gradient the first forward
    alone gives."""
    import torch

    D, AD = dev
    from pydyn_amd import meteo

    met = meteo.Meteo(meteo.grid("small"), seed=3)
    window = AD.Window(met, 2)
    i, j, obs = AD.samples(met.g, n=40)
    init = clipped_init(met)

    def grads(second=None):
        x = [D.upload(v).requires_grad_(True) for v in init]
        e = D.upload(met.emis).requires_grad_(True)
        out = AD.transport(window, *x, e)
        if second is not None:  # another forward from other tracers, its graph left alone
            AD.transport(window, *[D.upload(v * second) for v in init], D.upload(met.emis))
        AD.loss(out[0], window.m_final, i, j, obs).backward()
        return [D.download(v.grad) for v in (*x, e)]

    alone = grads()
    interleaved = grads(second=1.5)
    for a, b in zip(alone, interleaved):
        assert np.array_equal(a.view(np.int64), b.view(np.int64))
    assert torch.cuda.is_available()


@pytest.fixture(scope="module")
def c4(dev):
    """The 120 x 90 x 34 grid, one tracer, 24 hourly intervals."""
    from pydyn_amd import meteo

    met = meteo.Meteo(meteo.grid("glb3x2"), seed=1)
    return met


def test_dot_product_24h_glb3x2(dev, c4):
    """Limiter off the tracer map is linear in (initial tracers, emission): the forward run from
    (d, e) is the tangent, and <F(d, e), l> = <d, ADM_d l> + <e, ADM_e l> to 1e-12."""
    D, AD = dev
    print("dot product 24 h glb3x2", flush=True)
    met = c4
    window = AD.Window(met, 24, limits=False)
    rng = np.random.default_rng(15)
    st = met.state()
    shape = st["rm"].shape
    d = [st["rm"] * rng.standard_normal(shape) for _ in range(4)]
    e = met.emis * rng.standard_normal(met.emis.shape)
    lam = [rng.standard_normal(shape) for _ in range(4)]
    out = window.forward([D.upload(x) for x in d], D.upload(e))
    ad, ae = window.backward([D.upload(x) for x in lam])
    lhs = sum(float(np.sum(D.download(a) * b)) for a, b in zip(out, lam))
    rhs = sum(float(np.sum(a * D.download(b))) for a, b in zip(d, ad)) + float(np.sum(e * D.download(ae)))
    print(f"dot product: {lhs!r} {rhs!r} rel {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_dot_product_limiter_on_24h_glb3x2_oracle_tangent(dev, c4):
    """Limiter on, from slopes the limiter clips: the NumPy tangent-linear model run alongside
    the oracle's forward over 24 hourly intervals at 120 x 90 x 34 against the device adjoint of
    the same (bit-identical) trajectory: <TLM (d, e), l> = <d, ADM_d l> + <e, ADM_e l> to 1e-12."""
    from oracle import adjoint as A

    D, AD = dev
    met = c4
    init = clipped_init(met)
    rng = np.random.default_rng(18)
    shape = init[0].shape
    scale = float(np.abs(init[0]).mean())
    dx = [1e-3 * scale * rng.standard_normal(shape) for _ in range(4)]
    de = 1e-3 * float(np.abs(met.emis).mean()) * rng.standard_normal(met.emis.shape)
    lam = [rng.standard_normal(shape) for _ in range(4)]
    print("oracle tangent over 24 h of glb3x2 (a few minutes of CPU)", flush=True)
    tl, w = A.Window.tangent(met, 24, dx, de, init=init, progress=lambda k: print(f"  interval {k}", flush=True))
    window = AD.Window(met, 24)
    final = window.forward([D.upload(x) for x in init], D.upload(met.emis))
    for k, (g, f) in enumerate(zip(final, w.final)):
        assert np.array_equal(D.download(g), f), f"forward {FIELDS[k]}"
    ad, ae = window.backward([D.upload(x) for x in lam])
    lhs = sum(float(np.sum(a * b)) for a, b in zip(tl, lam))
    rhs = sum(float(np.sum(a * D.download(b))) for a, b in zip(dx, ad)) + float(np.sum(de * D.download(ae)))
    print(f"dot product, limiter on: {lhs!r} {rhs!r} rel {abs(lhs - rhs) / abs(lhs):.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gradient_24h_glb3x2_finite_differences(dev, c4):
    """The loss's gradient through the autograd Function against central differences of the
    forward model, limiter on, for the emission and the initial state."""
    import torch

    D, AD = dev
    print("finite differences 24 h glb3x2", flush=True)
    met = c4
    window = AD.Window(met, 24)
    i, j, obs = AD.samples(met.g)
    st = met.state()
    init = [D.upload(st[k]) for k in FIELDS]
    emis = D.upload(met.emis)
    x = [t.clone().requires_grad_(True) for t in init]
    e = emis.clone().requires_grad_(True)
    out = AD.transport(window, *x, e)
    L = AD.loss(out[0], window.m_final, i, j, obs)
    L.backward()
    g_e = D.download(e.grad)
    g_x = [D.download(t.grad) for t in x]
    assert np.abs(g_e).max() > 0.0 and np.abs(g_x[0]).max() > 0.0
    rng = np.random.default_rng(16)
    de = met.emis * rng.standard_normal(met.emis.shape)
    dx = [st["rm"] * rng.standard_normal(st["rm"].shape) for _ in range(4)]

    def loss_at(eps_e, eps_x):
        with torch.no_grad():
            xs = [t + D.upload(eps_x * d) for t, d in zip(init, dx)]
            out = window.forward(xs, emis + D.upload(eps_e * de))
            return float(AD.loss(out[0], window.m_final, i, j, obs))

    for what, eps_e, eps_x, pred in (
        ("emission", 1e-3, 0.0, float(np.sum(g_e * de))),
        ("initial state", 0.0, 1e-3, sum(float(np.sum(g * d)) for g, d in zip(g_x, dx))),
    ):
        fd = (loss_at(eps_e, eps_x) - loss_at(-eps_e, -eps_x)) / 2.0
        err = abs(fd - pred * max(eps_e, eps_x)) / abs(pred * max(eps_e, eps_x))
        print(f"{what}: fd {fd!r} adjoint {pred * max(eps_e, eps_x)!r} rel {err:.2e}", flush=True)
        assert err <= 1e-6, what
