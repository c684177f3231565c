"""The oracle's Calc_Kzz (test infrastructure): pydyn/diffusion.py:71-276 restated level by level
over all columns at once -- the friction velocity (dd_calc_ustar), the interface heights
(Meteo_Produce_Gph), the step's air mass and winds, bldiff's free-troposphere and boundary-layer
diffusivities with the two bulk-Richardson sweeps for the boundary-layer height, and dkg.

Every `if` of the Fortran is the 0/1 mask jqif(x, y) multiplying both branches, so a NaN or an
inf in an unused branch reaches the result as it does there; max / min propagate NaN as NumPy's
do.  Products and quotients are written in the reference's left-to-right order; log, pow and exp
are the oracle's fdlibm (oracle/libm.c), cos (for the zonal cell widths) the C library's.

    ustar         pydyn/diffusion.py:76-91
    geopotential  pydyn/diffusion.py:94-106
    mass_t, winds pydyn/diffusion.py:109-124
    bldiff        pydyn/diffusion.py:127-253 (the sweeps :127-155)
    dkg           pydyn/diffusion.py:256-263
    calc_kzz      pydyn/diffusion.py:266-276
"""

import math

import numpy as np

from . import libm

GRAV, PI = 9.80665, 3.14159265358979
AE, VKARMAN, CP_AIR, LVAP, RGAS, XMAIR = 6.371e6, 0.4, 1004.0, 2.5e6, 8.3144, 28.94
GTOR = PI / 180.0
RHO_AIR = 1.2
CKMIN, SFFRAC, ONET = 1.0e-15, 0.1, 1.0 / 3.0
RAIR = RGAS * 1.0e3 / XMAIR
EPSILO = RAIR / 461.51
APZERO = 100000.0
CCON = SFFRAC * VKARMAN
BETAH, BINM, FAK, FAKN = 15.0, 15.0 * SFFRAC, 8.5, 7.2
ZKAPPA = RAIR / CP_AIR
ZCRDQ = 1.0 / EPSILO - 1.0
TINY, ZACB, RICR, SSFRAC, KVHMIN = 1.0e-9, 100.0, 0.3, 0.1, 0.1


def jqif(x, y):
    """1.0 where x - y is +0.0 or positive (its sign bit clear), else 0.0."""
    return np.where(np.signbit(np.asarray(x, dtype=np.float64) - y), 0.0, 1.0)


def _max(a, b):
    return np.maximum(a, b)


def _min(a, b):
    return np.minimum(a, b)


def ustar(lsm, sr, srols, u10m, v10m, ewss, nsss):
    xland = lsm / 100.0
    tau = np.sqrt(ewss * ewss + nsss * nsss)
    ustar_sea = _max(0.001, np.sqrt(tau / RHO_AIR))
    wind10m = np.sqrt(u10m * u10m + v10m * v10m)
    with np.errstate(all="ignore"):
        sr_land = _max(1e-2, srols)
        sr_help = _min(sr, 0.03)
        land = ((VKARMAN * wind10m) / libm.log(10.0 / sr_help)) * libm.log(75.0 / sr_help)
        land = land / libm.log(10.0 / sr_land)
    ustar_land = np.where(xland > 0.0, land, 0.0)
    return xland * ustar_land + (1.0 - xland) * ustar_sea


def geopotential(oro, sp, t, q, at, bt):
    im, jm, lm = t.shape
    gph = np.empty((im, jm, lm + 1))
    gph[..., 0] = oro / GRAV
    for k in range(lm - 1):
        tv = t[..., k] * (1.0 + 0.608 * q[..., k])
        ratio = (at[k] + bt[k] * sp) / (at[k + 1] + bt[k + 1] * sp)
        gph[..., k + 1] = gph[..., k] + ((tv * 287.05) * libm.log(ratio)) / GRAV
    gph[..., lm] = 200000.0
    return gph


def _sweep(pblh, whm, uwind, vwind, wthm, q, ustr, wtseff, reference):
    """A bulk-Richardson sweep up the columns (reference None: the first, against wtseff)."""
    lm = whm.shape[-1]
    first = reference is None
    jwork, wrino = np.ones_like(pblh), np.zeros_like(pblh)
    with np.errstate(all="ignore"):
        for k in range(0 if first else 1, lm):
            du, dv = uwind[..., k] - uwind[..., 0], vwind[..., k] - vwind[..., 0]
            vvk = _max((du * du + dv * dv) + (ZACB * ustr) * ustr, TINY)
            tkv = wthm[..., k] * (1.0 + ZCRDQ * q[..., k])
            dz = whm[..., k] - whm[..., 0]
            if first:
                zrino = ((GRAV * (tkv - wtseff)) * dz) / (wtseff * vvk)
            else:
                zrino = ((GRAV * (tkv - reference)) * dz) / (vvk * wtseff)
            zrino = jwork * (zrino + np.copysign(1.0e-10, zrino))
            jq = jqif(RICR, zrino)
            frac = (RICR - wrino) / (zrino - wrino)
            at_ricr = frac * whm[..., 0] if k == 0 else whm[..., k - 1] + frac * (whm[..., k] - whm[..., k - 1])
            pblh = jq * pblh + (1.0 - jq) * at_ricr
            jwork = jq * jwork
            wrino = zrino + (1.0 - jwork) * 0.1
    top = whm[..., lm - 1]
    return (1.0 - jwork) * pblh + (top * jwork if first else jwork * top)


def bldiff(phlb, gph, t, q, m, pu, pv, ustar_loc, sshf, slhf, dxx, dyy):
    im, jm, lm = t.shape
    lo, hi = phlb[..., :-1], phlb[..., 1:]
    z0 = gph[..., :1]
    with np.errstate(all="ignore"):
        pf = (lo + hi) * 0.5
        wthm = t * libm.pow(APZERO / pf, ZKAPPA)
        zthvk = wthm * (1.0 + ZCRDQ * q)
        intfac = (lo - pf) / (lo - hi)
        whm = (gph[..., :-1] - z0) * (1.0 - intfac) + (gph[..., 1:] - z0) * intfac
        dz1, dp1 = gph[..., 1] - gph[..., 0], phlb[..., 0] - phlb[..., 1]
        wheat = ((((-sshf) * dz1) * GRAV) / dp1) / CP_AIR
        wqflx = ((((-slhf) * dz1) * GRAV) / dp1) / LVAP
        ustr = _max(ustar_loc, 0.01)
        uwind = (((dxx[None, :, None] * (pu[1:] + pu[:-1])) * 0.5) / m)
        vwind = ((dyy * (pv[:, 1:] + pv[:, :-1])) * 0.5) / m

        z = gph[..., 1:lm] - z0
        jq = jqif(1000.0, z)
        zlambdac = jq * 300.0 + (1.0 - jq) * (30.0 + 270.0 * libm.exp(1.0 - z / 1000.0))
        cml2 = 1.0 / (1.0 / zlambdac + 1.0 / (VKARMAN * z))
        cml2 = cml2 * cml2
        du, dv = uwind[..., 1:] - uwind[..., :-1], vwind[..., 1:] - vwind[..., :-1]
        zdz = whm[..., 1:] - whm[..., :-1]
        zdup2 = _max(du * du + dv * dv, 1.0e-10) / (zdz * zdz)
        log_pf = libm.log(pf)
        arg = (libm.log(phlb[..., 1:lm]) - log_pf[..., :-1]) / (log_pf[..., 1:] - log_pf[..., :-1])
        dthv = zthvk[..., 1:] - zthvk[..., :-1]
        zthva = zthvk[..., :-1] + arg * dthv
        zrinub = (((GRAV / zthva) * dthv) / zdz) / zdup2
        zfunst = np.sqrt(_max(1.0 - 18.0 * zrinub, 0.0))
        zfstabh = 1.0 / (1.0 + (10.0 * zrinub) * (1.0 + 8.0 * zrinub))
        zkvn = cml2 * np.sqrt(zdup2)
        jq = jqif(zrinub, 0.0)
        wkvf = jq * _max(CKMIN, zkvn * zfstabh) + (1.0 - jq) * _max(CKMIN, zkvn * zfunst)

        wtseff = wthm[..., 0] * (1.0 + ZCRDQ * q[..., 0])
        wheatv = wheat + (ZCRDQ * wthm[..., 0]) * wqflx
        cube = (ustr * ustr) * ustr
        wobkl = ((-wtseff) * cube) / ((GRAV * VKARMAN) * (wheatv + np.copysign(1.0e-10, wheatv)))
        pblh = _sweep(np.zeros((im, jm)), whm, uwind, vwind, wthm, q, ustr, wtseff, None)
        jq = jqif(wheatv, 0.0)
        fmt = libm.pow(jq * (1.0 - (BINM * pblh) / wobkl) + (1.0 - jq), ONET)
        zexcess = (wheatv * FAK) / (ustr * fmt)
        thvparc = wtseff + jq * zexcess
        pblh = _max(_sweep(pblh, whm, uwind, vwind, wthm, q, ustr, wtseff, thvparc), 100.0)

        P, W, HV, U, TS = (a[..., None] for a in (pblh, wobkl, wheatv, ustr, wtseff))
        jq = jqif(P, z)
        zh = (jq * z) / P
        zl = (jq * z) / W + (1.0 - jq)
        zzh = jq * ((1.0 - zh) * (1.0 - zh))
        jq1 = jqif(HV, 0.0)
        jq2 = jq * (1.0 - jq1)
        jq3 = jq * jq1
        jq6 = (jq1 * jqif(P, whm[..., :-1])) * jqif(whm[..., 1:], P)
        jq1 = jqif(SFFRAC, zh)
        jq4 = jq3 * jq1
        jq5 = jq3 * (1.0 - jq1)
        cml2 = 1.0 / (1.0 / (VKARMAN * z) + 1.0 / 450.0)
        cml2 = cml2 * cml2
        jqq = jqif(zrinub, 0.0)
        zfstabh = jqq / (1.0 + (10.0 * zrinub) * np.sqrt(1.0 + jqq * zrinub)) + (1.0 - jqq)
        zpblk = (cml2 * np.sqrt(zdup2)) * zfstabh
        zkvh = jq2 * _max(zpblk, KVHMIN) + (1.0 - jq2) * wkvf
        zslask = jq4 * (1.0 - BETAH * zl) + (1.0 - jq4)
        zpblk = (1.0 - jq4) * zkvh + jq4 * ((((U * VKARMAN) * z) * zzh) * libm.pow(zslask, ONET))
        zslask = jq5 * (1.0 - ((SSFRAC * BETAH) * P) / W) + (1.0 - jq5)
        wsc = U * libm.pow(zslask, ONET)
        zpblk = (1.0 - jq5) * zpblk + (((jq5 * wsc) * VKARMAN) * z) * zzh
        zwstr = libm.pow(((np.abs(HV) * GRAV) * P) / TS, ONET)
        obukov = jq3 * W - (1.0 - jq3)
        stab = 1.0 - ((SSFRAC * BETAH) * P) / obukov
        term1 = libm.pow(stab, ONET)
        term3 = ((CCON * FAKN) * zwstr) / (U * term1)
        pr = jq3 * (term1 / np.sqrt(stab) + term3) + (1.0 - jq3)
        kvh = jq3 * _max(zpblk / pr, KVHMIN) + (1.0 - jq3) * zkvh
        kvhentr = (0.2 * HV) / (dthv / zdz)
        kvh = jq6 * kvhentr + (1.0 - jq6) * kvh
    return kvh, pblh


def dkg(kvh, m, gph, lmax_conv):
    n = lmax_conv - 1
    out = np.zeros((*m.shape[:2], lmax_conv))
    dz = gph[..., 2 : n + 2] - gph[..., :n]
    out[..., :n] = ((_max(0.0, kvh[..., :n]) * 2.0) * (m[..., 1 : n + 1] + m[..., :n])) / (dz * dz)
    return out


def zonal_widths(jm, dx, dy, ybeg):
    """dxx (jm): the cells' zonal width in m at their centre latitudes, cos from the C library."""
    lat = (ybeg + 0.5 * dy) + dy * np.arange(jm)
    c = np.array([math.cos(v) for v in lat * GTOR])
    return ((AE * dx) * GTOR) * c


def calc_kzz(spm, sp, pu, pv, t, q, surface, at, bt, dxyp, lmax_conv, dx=6.0, dy=4.0, ybeg=-90.0):
    ustar_loc = ustar(*(surface[k] for k in ("lsm", "sr", "srols", "u10m", "v10m", "ewss", "nsss")))
    gph = geopotential(surface["oro"], sp, t, q, at, bt)
    phlb = at + bt * spm[..., None]
    m = ((phlb[..., :-1] - phlb[..., 1:]) * dxyp[None, :, None]) / GRAV
    dxx = zonal_widths(t.shape[1], dx, dy, ybeg)
    dyy = (AE * dy) * GTOR
    kvh, pblh = bldiff(phlb, gph, t, q, m, pu, pv, ustar_loc, surface["sshf"], surface["slhf"], dxx, dyy)
    return dkg(kvh, m, gph, lmax_conv), {"ustar": ustar_loc, "gph": gph, "pblh": pblh, "kvh": kvh}
