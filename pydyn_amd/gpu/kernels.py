"""The device operator layer: the reference's pydyn.gpu.kernels names and argument orders on
device arrays (pydyn_amd.device), each a call into libtm5hip.

Operators are asynchronous on the context's stream and in place, as the reference's are.  The
advection operators check the device flags before returning unless `deferred=True` (then
`device.context().raise_errors()` reports them later), raising the reference's CFLError /
ValueError with its messages.
"""

from .. import _lib
from .. import device as D
from ..device import CFLError  # noqa: F401  (the reference's exception, re-exported)

__all__ = [
    "CFLError",
    "add_2d",
    "advect_x",
    "advect_x_rows",
    "advect_y",
    "advect_z",
    "convec",
    "decay",
    "dkg",
    "fluxes_from_meteo",
    "m2phlb1",
    "mass_x_rows",
    "mass_y",
    "mass_z",
    "scale",
]


def _dims(m):
    im, jm, lm = m.shape
    return im, jm, lm


def _tr(rm, rxm, rym, rzm):
    if rm.dim() == 3:
        rm, rxm, rym, rzm = (q.unsqueeze(3) for q in (rm, rxm, rym, rzm))
    return D.tracers(rm, rxm, rym, rzm)


def _finish(ctx, deferred):
    if not deferred:
        ctx.raise_errors()


# ------------------------------------------------------------------ advection
def advect_x(m, rm, rxm, rym, rzm, am, limits=True, max_nloop=50, deferred=False):
    """dynamu on every (j, l) row (pydyn/gpu/kernels.py:64, pydyn/kernels.py:129-219).  Returns
    the largest nloop of the rows."""
    ctx = D.context()
    im, jm, lm = _dims(m)
    _lib.check(_lib.lib().tm5_advect_x(ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm), D.view(am),
                                       int(limits), int(max_nloop)), "tm5_advect_x")  # fmt: skip
    _finish(ctx, deferred)
    return None if deferred else ctx.nloop(0)[0]


def advect_x_rows(m, rm, rxm, rym, rzm, a, nrow=None, limits=True, max_nloop=None, deferred=False):
    """pydyn/gpu/kernels.py:64: advect_x on uniform rows (a reduced grid's nrow is not supported)."""
    if nrow is not None and any(int(n) != m.shape[0] for n in nrow):
        raise NotImplementedError("advect_x_rows: reduced rows are not supported")
    return advect_x(m, rm, rxm, rym, rzm, a, limits, 50 if max_nloop is None else max_nloop, deferred)


def advect_y(m, rm, rxm, rym, rzm, bm, limits=True, max_nloop=10, stats=None, deferred=False):
    """dynamv, one nloop per layer (pydyn/gpu/kernels.py:103, pydyn/kernels.py:223-291).
    Returns the per-layer nloop."""
    ctx = D.context()
    im, jm, lm = _dims(m)
    _lib.check(_lib.lib().tm5_advect_y(ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm), D.view(bm),
                                       int(limits), int(max_nloop)), "tm5_advect_y")  # fmt: skip
    _finish(ctx, deferred)
    if deferred:
        return None
    layers = ctx.nloop(lm)[1]
    if stats is not None:
        stats["nloop_max"] = int(layers.max())
    return layers


def advect_z(m, rm, rxm, rym, rzm, cm, limits=True, iter_limit=15, stats=None, deferred=False):
    """dynamw, one n_advz for the region (pydyn/gpu/kernels.py:227, pydyn/kernels.py:295-346).
    Returns n_advz."""
    ctx = D.context()
    im, jm, lm = _dims(m)
    _lib.check(_lib.lib().tm5_advect_z(ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm), D.view(cm),
                                       int(limits), int(iter_limit)), "tm5_advect_z")  # fmt: skip
    _finish(ctx, deferred)
    if deferred:
        return None
    n = ctx.nloop(0)[2]
    if stats is not None:
        stats["nloop_max"] = n
    return n


# ------------------------------------------------------------------ Check_CFL's mass-only kernels
def _mass(fn, m, flux, cap):
    ctx = D.context()
    ctx.flags(clear=True)
    im, jm, lm = _dims(m)
    _lib.check(getattr(_lib.lib(), fn)(ctx.ptr, im, jm, lm, D.view(m), D.view(flux), int(cap)), fn)
    f = ctx.flags(clear=True)
    if f & _lib.ERRORS:
        raise ValueError(f"{fn}: device error flags {f}")
    return not (f & _lib.REHEARSAL)


def mass_x_rows(m, a, nrow=None, max_nloop=10):
    """dynamum (pydyn/gpu/kernels.py:420, pydyn/cfl.py:140-167): True when every row stayed
    under its cap (m, a updated), False otherwise."""
    if nrow is not None and any(int(n) != m.shape[0] for n in nrow):
        raise NotImplementedError("mass_x_rows: reduced rows are not supported")
    return _mass("tm5_mass_x", m, a, max_nloop)


def mass_y(m, bm, max_nloop=1):
    """dynamvm (pydyn/gpu/kernels.py:447, pydyn/cfl.py:170-183)."""
    return _mass("tm5_mass_y", m, bm, max_nloop)


def mass_z(m, cm, iter_limit=1):
    """dynamwm (pydyn/gpu/kernels.py:497, pydyn/zoom/check_cfl.py:206-216)."""
    return _mass("tm5_mass_z", m, cm, iter_limit)


# ------------------------------------------------------------------ mass fluxes
def fluxes_from_meteo(pu, pv, bt, dt, am, bm, cm, amplitude=1.0):
    """dynam0 into am, bm, cm (pydyn/gpu/kernels.py:300 with pydyn/zoom/massflow.py:446-461):
    bt is a device array (lm+1,)."""
    ctx = D.context()
    im, jm, lm = pu.shape[0] - 1, pu.shape[1], pu.shape[2]
    _lib.check(_lib.lib().tm5_dynam0(ctx.ptr, im, jm, lm, D.view(pu), D.view(pv), float(amplitude), bt.data_ptr(),
                                     float(dt), D.view(am), D.view(bm), D.view(cm)), "tm5_dynam0")  # fmt: skip


def m2phlb1(m, dxyp, at, bt, sp, phlb):
    """sp and phlb from the column mass (pydyn/zoom/massflow.py:282-290)."""
    ctx = D.context()
    im, jm, lm = _dims(m)
    _lib.check(_lib.lib().tm5_m2phlb1(ctx.ptr, im, jm, lm, D.view(m), dxyp.data_ptr(), at.data_ptr(), bt.data_ptr(),
                                      D.view(sp), D.view(phlb)), "tm5_m2phlb1")  # fmt: skip


# ------------------------------------------------------------------ physics
def convec(m, rm, rxm, rym, rzm, conv, dt, lmax):
    """convec on every column (pydyn/gpu/kernels.py:351, pydyn/kernels.py:423-457): conv holds
    device arrays entu, detu, entd, detd, dkg (float64) and cloud_top, cloud_lfs (int32)."""
    ctx = D.context()
    im, jm, lm = _dims(m)
    _lib.check(_lib.lib().tm5_convec(
        ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm),
        *(D.view(conv[k]) for k in ("entu", "detu", "entd", "detd")),
        D.iview(conv["cloud_top"]), D.iview(conv["cloud_lfs"]), D.view(conv["dkg"]), float(dt), int(lmax),
    ), "tm5_convec")  # fmt: skip


def dkg(kvh, m, gph, lmax):
    """dkg (im, jm, lmax) from kvh, m, gph (pydyn/diffusion.py:256-263)."""
    ctx = D.context()
    im, jm, _ = _dims(m)
    out = D.empty((im, jm, lmax))
    _lib.check(_lib.lib().tm5_dkg(ctx.ptr, im, jm, D.view(kvh), D.view(m), D.view(gph), int(lmax), D.view(out)),
               "tm5_dkg")  # fmt: skip
    return out


def add_2d(emis, factor, ratio, efrac, rm, rxm, rym, rzm, level=1, tracer=1):
    """do_add_2d of inject(emis, factor, ...) (pydyn/emis.py:52-62, pydyn/zoom/physics.py:516-525)."""
    ctx = D.context()
    im, jm = emis.shape[:2]
    _lib.check(_lib.lib().tm5_add_2d(ctx.ptr, im, jm, D.view(emis), float(factor), float(ratio), float(efrac),
                                     _tr(rm, rxm, rym, rzm), int(level), int(tracer)), "tm5_add_2d")  # fmt: skip


def decay(loss, dt, rm, rxm, rym, rzm, tracers=None):
    """The CH4 decay of the 1-based `tracers` (all by default) over dt (pydyn/chem.py:168-178,
    pydyn/zoom/physics.py:528-539)."""
    ctx = D.context()
    im, jm, lm = loss.shape[:3]
    q = _tr(rm, rxm, rym, rzm)
    tr = list(range(1, q.nt + 1)) if tracers is None else sorted(int(t) for t in tracers)
    if tr and tr != list(range(tr[0], tr[-1] + 1)):
        raise ValueError("decay: the tracers must be consecutive")
    if tr:
        _lib.check(_lib.lib().tm5_decay(ctx.ptr, im, jm, lm, D.view(loss), float(dt), q, tr[0] - 1, len(tr)),
                   "tm5_decay")  # fmt: skip


def scale(a, factor):
    """a = a * factor (pydyn/zoom/host.py:226-230)."""
    ctx = D.context()
    n0, n1, n2 = a.shape
    _lib.check(_lib.lib().tm5_scale(ctx.ptr, n0, n1, n2, D.view(a), float(factor)), "tm5_scale")
