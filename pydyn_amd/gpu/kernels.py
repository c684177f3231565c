"""The device operator layer: the reference's pydyn.gpu.kernels names and argument orders on
device arrays (pydyn_amd.device), each a call into libtm5hip.

Operators are asynchronous on the context's stream and in place, as the reference's are.  The
advection operators check the device flags before returning unless `deferred=True` (then
`device.context().raise_errors()` reports them later), raising the reference's CFLError /
ValueError with its messages.
"""

import numpy as np

from .. import _lib
from .. import device as D
from ..device import CFLError  # noqa: F401  (the reference's exception, re-exported)

__all__ = [
    "CFLError",
    "ok_flag",
    "add_2d",
    "advect_x",
    "advect_x_rows",
    "advect_y",
    "advect_z",
    "advect_z_substeps",
    "cluster_sum",
    "convec",
    "decay",
    "dkg",
    "fluxes_from_meteo",
    "loss_rates",
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
    """dynamu on every (j, l) row of a region at once (pydyn/gpu/kernels.py:64): m (n, jm, lm),
    tracers (n, jm, lm, nt), a (n+1, jm, lm) with a[0] == a[n] per row, nrow (jm,) int32 device
    array of the cells of each latitude (a reduced row's cells first; None: every row whole).
    Returns the device arrays (nloop, alpha_max), one value per row j + jm l."""
    torch = D._torch()
    ctx = D.context()
    im, jm, lm = _dims(m)
    nloop = torch.empty(jm * lm, dtype=torch.int32, device="cuda")
    amax = torch.empty(jm * lm, dtype=torch.float64, device="cuda")
    rows = None if nrow is None else nrow.to(torch.int32).contiguous()
    _lib.check(_lib.lib().tm5_advect_x_rows(ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm), D.view(a),
                                            None if rows is None else rows.data_ptr(), int(limits),
                                            50 if max_nloop is None else int(max_nloop), nloop.data_ptr(),
                                            amax.data_ptr()), "tm5_advect_x_rows")  # fmt: skip
    _finish(ctx, deferred)
    return nloop, amax


def _layer_halo(full, bm, layers, level_offset):
    """dynamv's whole-layer scaling of `full` outside its view bm (the reference's layer scale,
    pydyn/gpu/csrc/advect.cuh:192-205 and pydyn/zoom/advect.py:212-226): per layer that
    sub-cycled, the faces outside bm through the retry scalings (n-1)/n, n = 2..nloop, and back
    times nloop.  bm is a slice of full with its strides."""
    torch = D._torch()
    off = bm.storage_offset() - full.storage_offset()
    origin = []
    for ax in sorted(range(3), key=lambda k: -full.stride(k)):
        origin.append((ax, off // full.stride(ax)))
        off %= full.stride(ax)
    start = dict(origin)  # the view's first element in full, per axis
    inside = torch.zeros(full.shape[:2], dtype=torch.bool, device=full.device)
    inside[start[0] : start[0] + bm.shape[0], start[1] : start[1] + bm.shape[1]] = True
    for lev, n in enumerate(layers):
        n = int(n)
        if n == 1:
            continue
        layer = full[:, :, lev + level_offset]
        halo = layer[~inside]
        for r in range(2, n + 1):
            halo = halo * float(r - 1) / float(r)
        layer[~inside] = halo * float(n)


def advect_y(m, rm, rxm, rym, rzm, bm, limits=True, max_nloop=10, stats=None, full_bm=None, level_offset=0,
             deferred=False):
    """dynamv, one nloop per layer (pydyn/gpu/kernels.py:103, pydyn/kernels.py:223-291).
    full_bm: the array whose whole layers dynamv scales and restores (region 1's halo'd bm, bm a
    slice of it), its levels offset by level_offset from bm's.  Returns the device array of the
    layers' nloop; stats gets nloop_max (its 'xi' is not computed here: None)."""
    torch = D._torch()
    ctx = D.context()
    im, jm, lm = _dims(m)
    _lib.check(_lib.lib().tm5_advect_y(ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm), D.view(bm),
                                       int(limits), int(max_nloop)), "tm5_advect_y")  # fmt: skip
    _finish(ctx, deferred)
    if deferred:
        return None
    layers = ctx.nloop(lm)[1]
    if full_bm is not None and full_bm is not bm:
        _layer_halo(full_bm, bm, layers, level_offset)
    if stats is not None:
        stats["nloop_max"], stats["xi"] = int(layers.max()), None
    return torch.as_tensor(layers, dtype=torch.int32).to("cuda")


def advect_z(m, rm, rxm, rym, rzm, cm, limits=True, iter_limit=15, stats=None, deferred=False):
    """dynamw, one n_advz for the region (pydyn/gpu/kernels.py:227, pydyn/kernels.py:295-346).
    Returns n_advz; stats gets nloop_max (its 'xi': None, as the reference's)."""
    ctx = D.context()
    im, jm, lm = _dims(m)
    _lib.check(_lib.lib().tm5_advect_z(ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm), D.view(cm),
                                       int(limits), int(iter_limit)), "tm5_advect_z")  # fmt: skip
    _finish(ctx, deferred)
    if deferred:
        return None
    n = ctx.nloop(0)[2]
    if stats is not None:
        stats["nloop_max"], stats["xi"] = n, None
    return n


def advect_z_substeps(m, rm, rxm, rym, rzm, cm, n_advz, limits=True, prelimit=False):
    """dynamw's n_advz sub-steps with cm / n_advz, cm not written (pydyn/gpu/kernels.py:262,
    pydyn/kernels.py:324-346); prelimit applies dynamw's limiter to rzm first.  n_advz: an int
    or a one-element device array."""
    ctx = D.context()
    im, jm, lm = _dims(m)
    n = int(n_advz.reshape(-1)[0]) if hasattr(n_advz, "reshape") else int(n_advz)
    _lib.check(_lib.lib().tm5_advect_z_substeps(ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm), D.view(cm),
                                                n, int(limits), int(prelimit)), "tm5_advect_z_substeps")  # fmt: skip


def cluster_sum(q, n, c):
    """seqsum(clusters(q, n, c), axis=1) of a (cells, a[, b]) device array: a new (n, a[, b])
    array, each cluster summed from its first cell (pydyn/gpu/kernels.py:327, pydyn/redgrid.py)."""
    ctx = D.context()
    v = q if q.dim() == 3 else q.unsqueeze(2)
    if v.shape[0] < n * c:
        raise ValueError("cluster_sum: fewer cells than n * c")
    out = D.empty((n, v.shape[1], v.shape[2]))
    _lib.check(_lib.lib().tm5_cluster_sum(ctx.ptr, int(n), int(c), v.shape[1], v.shape[2], D.view(v), D.view(out)),
               "tm5_cluster_sum")  # fmt: skip
    return out if q.dim() == 3 else out[..., 0]


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
    """dynamum (pydyn/gpu/kernels.py:420, pydyn/cfl.py:140-167) on rows of nrow[j] cells (None:
    every row whole): True when every row stayed under its cap (m, a updated), False otherwise."""
    torch = D._torch()
    ctx = D.context()
    ctx.flags(clear=True)
    im, jm, lm = _dims(m)
    rows = None if nrow is None else nrow.to(torch.int32).contiguous()
    _lib.check(_lib.lib().tm5_mass_x_rows(ctx.ptr, im, jm, lm, D.view(m), D.view(a),
                                          None if rows is None else rows.data_ptr(), int(max_nloop)),
               "tm5_mass_x_rows")  # fmt: skip
    f = ctx.flags(clear=True)
    if f & _lib.ERRORS:
        raise ValueError(f"tm5_mass_x_rows: device error flags {f}")
    return not (f & _lib.REHEARSAL)


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
def convec(m, rm, rxm, rym, rzm, conv, dt, lmax, cols=None, inverse="crout"):
    """convec (pydyn/gpu/kernels.py:351, pydyn/kernels.py:423-457): conv holds device arrays
    entu, detu, entd, detd, dkg (float64) and cloud_top, cloud_lfs (int32).  cols: the columns
    to act on (device int32, flat i * jm + j; None: every column).  inverse: 'crout' (fastminv,
    the reference's default) or 'lapack' (reference LAPACK's DGETRF + DGETRI per column, the
    with_lapack build; a singular matrix raises ValueError)."""
    if inverse not in ("crout", "lapack"):
        raise ValueError(f"pydyn.gpu.convec: inverse {inverse!r}: 'crout' or 'lapack' (an open one)")
    torch = D._torch()
    ctx = D.context()
    im, jm, lm = _dims(m)
    args = (ctx.ptr, im, jm, lm, D.view(m), _tr(rm, rxm, rym, rzm),
            *(D.view(conv[k]) for k in ("entu", "detu", "entd", "detd")),
            D.iview(conv["cloud_top"]), D.iview(conv["cloud_lfs"]), D.view(conv["dkg"]), float(dt), int(lmax))  # fmt: skip
    mask = None
    if cols is not None:
        flat = cols.to(torch.int64).reshape(-1)
        mask = torch.zeros(jm, im, dtype=torch.int32, device="cuda")  # (j, i): i fastest
        mask[flat % jm, flat // jm] = 1
    if inverse == "lapack":
        _lib.check(_lib.lib().tm5_convec_lapack(*args, None if mask is None else mask.data_ptr()), "tm5_convec_lapack")
        ctx.raise_errors()  # a zero pivot (the reference checks its info flag here)
        return
    if mask is None:
        _lib.check(_lib.lib().tm5_convec(*args), "tm5_convec")
        return
    _lib.check(_lib.lib().tm5_convec_masked(*args, mask.data_ptr()), "tm5_convec_masked")


def inverse_rows(A):
    """The LAPACK kind's inverse of each row-major n x n matrix of the device array A
    (ncol, n, n), C-contiguous: (Y, info), info a one-element device int32 array set to 1 when
    a pivot is exactly zero (pydyn/gpu/kernels.py:403-411)."""
    torch = D._torch()
    A = A.contiguous()
    ncol, n, _ = A.shape
    Y = torch.empty_like(A)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx = D.context()
    _lib.check(_lib.lib().tm5_inverse_rows(ctx.ptr, ncol, n, A.data_ptr(), Y.data_ptr(), info.data_ptr()),
               "tm5_inverse_rows")  # fmt: skip
    return Y, info


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


def loss_rates(dt_sec, temper, phlb, lat, strat_press, strat_oh, strat_cl, strat_od, oh_3d, pl_trop=None,
               cl_trop=None, soil=None, oh_scale=1.0, cl_scale=1.0, o1d_scale=1.0, soil_rate=0.0):  # fmt: skip
    """chem.loss_rates and chem.decay on the device (pydyn/gpu/kernels.py:545-645): (loss
    (im, jm, lm, 4), decay (im, jm, lm)).  Host arrays are uploaded, device ones used as they are;
    lat and strat_press stay host arrays (the tropopause and the zonal columns' edges are derived
    from them on the host, as the reference does)."""
    from .. import chem

    torch = D._torch()

    def dev(x):
        if x is None or isinstance(x, torch.Tensor):
            return x
        return D.upload(np.asarray(x, dtype=np.float64))

    def flat(x):  # a contiguous (jm, n) device copy of a zonal column set
        x = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, dtype=np.float64))
        return x.to(device="cuda", dtype=torch.float64).contiguous()

    edges = chem.strat_edges(np.asarray(strat_press, dtype=np.float64))
    if not (np.diff(edges, axis=-1) < 0.0).all():
        raise ValueError("rebin: input levels not strictly descending (the Fortran keeps the stale column)")
    t, p, oh = dev(temper), dev(phlb), dev(oh_3d)
    pl, cl, sl = dev(pl_trop), dev(cl_trop), dev(soil)
    if (pl is None) != (cl is None):
        raise ValueError("loss_rates: pl_trop and cl_trop go together")
    im, jm, lm = t.shape
    nstrat = edges.shape[-1] - 1
    ntrop = cl.shape[2] if cl is not None else 0
    trop = flat(chem.tropopause(np.asarray(lat, dtype=np.float64)))
    e, s_oh, s_cl, s_od = flat(edges), flat(strat_oh), flat(strat_cl), flat(strat_od)
    loss, dec = D.zeros((im, jm, lm, 4)), D.empty((im, jm, lm))
    none = None
    ctx = D.context()
    rc = _lib.lib().tm5_loss_rates(
        ctx.ptr, im, jm, lm, nstrat, ntrop, D.view(t), D.view(p), trop.data_ptr(), e.data_ptr(), s_oh.data_ptr(),
        s_cl.data_ptr(), s_od.data_ptr(), D.view(oh), D.view(pl) if pl is not None else none,
        D.view(cl) if cl is not None else none, D.view(sl) if sl is not None else none, float(oh_scale),
        float(cl_scale), float(o1d_scale), float(soil_rate), float(dt_sec), D.view(loss), D.view(dec))
    _lib.check(rc, "tm5_loss_rates")
    ctx.raise_errors()
    return loss, dec


def ok_flag(fail):
    """Check_CFL's verdict on a device fail flag (pydyn/gpu/kernels.py:216-223): True when the
    flag (a device int array whose first element a failed trial set) is zero, read now.  The
    reference also defers the read while it speculates; here a queued run verifies its trials
    on the device itself (tm5_tree_run, tm5_region_run)."""
    return not int(fail.reshape(-1)[0].item())
