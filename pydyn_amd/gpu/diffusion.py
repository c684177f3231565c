"""Calc_Kzz on the device: the reference's pydyn.gpu.diffusion.calc_kzz (gpu/diffusion.py:38)
over tm5_calc_kzz.  The zonal cell widths dxx are computed here on the host with the C library's
cos, as the reference does (its ops.libm cos); everything else on the device."""

import math

import numpy as np

from .. import _lib
from .. import device as D

__all__ = ["calc_kzz"]

AE, PI = 6.371e6, 3.14159265358979  # Binas: the Fortran's literals
GTOR = PI / 180.0
SURFACE = ("oro", "lsm", "sr", "srols", "u10m", "v10m", "sshf", "slhf", "ewss", "nsss")


def zonal_widths(jm, dx, dy, ybeg):
    """dxx (jm) in m: AE dx GTOR cos(lat GTOR) at the cells' centre latitudes."""
    lat = (ybeg + 0.5 * dy) + dy * np.arange(jm)
    return ((AE * dx) * GTOR) * np.array([math.cos(v) for v in lat * GTOR])


def calc_kzz(spm, sp, pu, pv, t, q, surface, at, bt, dxyp, lmax_conv, dx=6.0, dy=4.0, ybeg=-90.0, pblh=None,
             parts=None):  # fmt: skip
    """dkg (im, jm, lmax_conv) of one step as a device array.  `pblh` (a device (im, jm) array) is
    filled with the boundary-layer height when given; `parts` (a dict) receives the device
    intermediates ustar, gph, pblh and kvh when given."""
    torch = D._torch()

    def dev(x):
        return x if isinstance(x, torch.Tensor) else D.upload(np.asarray(x, dtype=np.float64))

    def flat(x):
        x = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x, dtype=np.float64))
        return x.to(device="cuda", dtype=torch.float64).contiguous()

    t, q, pu, pv, spm, sp = (dev(x) for x in (t, q, pu, pv, spm, sp))
    im, jm, lm = t.shape
    surf = [dev(surface[k]) for k in SURFACE]
    views = (_lib.View * 10)(*(D.view(s) for s in surf))
    at_d, bt_d, dxyp_d = flat(at), flat(bt), flat(dxyp)
    dxx = flat(zonal_widths(jm, dx, dy, ybeg))
    dyy = (AE * dy) * GTOR
    gph, dkg = D.empty((im, jm, lm + 1)), D.empty((im, jm, int(lmax_conv)))
    ust, kvh = (D.empty((im, jm)), D.empty((im, jm, lm - 1))) if parts is not None else (None, None)
    if pblh is None and parts is not None:
        pblh = D.empty((im, jm))
    ctx = D.context()
    rc = _lib.lib().tm5_calc_kzz(
        ctx.ptr, im, jm, lm, int(lmax_conv), D.view(spm), D.view(sp), D.view(pu), D.view(pv), D.view(t), D.view(q),
        views, at_d.data_ptr(), bt_d.data_ptr(), dxyp_d.data_ptr(), dxx.data_ptr(), float(dyy), D.view(gph),
        D.view(dkg), D.view(ust) if ust is not None else None, D.view(pblh) if pblh is not None else None,
        D.view(kvh) if kvh is not None else None)  # fmt: skip
    _lib.check(rc, "tm5_calc_kzz")
    if parts is not None:
        parts.update({"ustar": ust, "gph": gph, "pblh": pblh, "kvh": kvh})
    return dkg
