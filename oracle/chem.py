"""The oracle's CH4 loss-rate construction (test infrastructure): pydyn/chem.py:59-178 restated
per column.  `rebin` follows the Fortran loop of rebin_levels(ip_mix = .true.) as the reference's
own test writes it (tests/test_physics.py:29-48: the j_start skip, the early exit, each layer's
overlaps summed from zero in order, the mean where the layer has thickness) -- the reference's
test asserts its vectorised rebin equals that loop bitwise; the golden vectors pin both here.

    strat_edges  pydyn/chem.py:106-113   the zonal columns' edges in Pa from pressures in hPa
    tropopause   pydyn/chem.py:116-119   30000 - 21500 cos^2(lat) Pa (cos of the C library)
    rebin        pydyn/chem.py:65-103    thickness-weighted layer means, descending edges
    loss_rates   pydyn/chem.py:129-160   OH (Spivakovsky below the tropopause), Cl (+ the
                                         tropospheric Cl rate), O(1D), soil at the surface
    decay        pydyn/chem.py:163-165   exp(-(((L1 + L2) + L3) + L4) dt)
exp and pow are the oracle's fdlibm (oracle/libm.c), the functions the reference runs with when
it is given that library (ops.use_libm).
"""

import math

import numpy as np

from . import libm

OH_A, OH_EA = 2.45e-12, 1775.0
CL_A, CL_T0, CL_N, CL_EA = 2.36e-12, 298.0, 1.37, 939.0
TROP_P0, TROP_DP = 30000.0, 21500.0


def strat_edges(press):
    """(…, n) hPa -> (…, n+1) edges in Pa: a half layer below the first, midpoints, 0 on top."""
    press = np.asarray(press, dtype=np.float64)
    n = press.shape[-1]
    edges = np.empty((*press.shape[:-1], n + 1))
    edges[..., 0] = press[..., 0] + 0.5 * (press[..., 0] - press[..., 1])
    edges[..., 1:n] = (press[..., :-1] + press[..., 1:]) / 2.0
    edges[..., n] = 0.0
    return edges * 100.0


def tropopause(lat):
    c = np.array([math.cos(v) for v in np.ravel(lat)]).reshape(np.shape(lat))
    return TROP_P0 - TROP_DP * (c * c)


def _descending(levels):
    if not (np.diff(levels) < 0.0).all():
        raise ValueError("rebin: levels not strictly descending (the Fortran keeps the stale column)")


def rebin(ip_levels, ip_data, op_levels):
    """One column: `ip_data` (nip) on the descending edges `ip_levels` (nip+1) averaged over the
    layers of the descending `op_levels` (nop+1), as the Fortran loops it."""
    _descending(ip_levels)
    _descending(op_levels)
    nip, nop = len(ip_data), len(op_levels) - 1
    out = np.zeros(nop)
    j_start = 0
    for i in range(nop):
        op_bot, op_top = op_levels[i], op_levels[i + 1]
        while j_start < nip and ip_levels[j_start + 1] >= op_bot:
            j_start += 1
        acc = 0.0
        for j in range(j_start, nip):
            ip_bot, ip_top = ip_levels[j], ip_levels[j + 1]
            if ip_bot <= op_top:
                break
            thickness = min(op_bot, ip_bot) - max(op_top, ip_top)
            if thickness > 0.0:
                acc = acc + ip_data[j] * thickness
        width = op_bot - op_top
        out[i] = acc / width if width > 0.0 else acc
    return out


def loss_rates(temper, phlb, lat, strat_press, strat_oh, strat_cl, strat_od, oh_3d, pl_trop=None, cl_trop=None,
               soil=None, oh_scale=1.0, cl_scale=1.0, o1d_scale=1.0, soil_rate=0.0):  # fmt: skip
    """The loss rates (im, jm, lm, 4), column by column."""
    im, jm, lm = temper.shape
    edges = strat_edges(strat_press)  # (jm, nstrat+1)
    trop = tropopause(lat)
    krate_oh = OH_A * libm.exp(-OH_EA / temper)
    oh_spiv = (krate_oh * oh_3d) * oh_scale
    if cl_trop is not None:
        krate_cl = (CL_A * libm.pow(temper / CL_T0, CL_N)) * libm.exp(-CL_EA / temper)
    loss = np.zeros((im, jm, lm, 4))
    for i in range(im):
        for j in range(jm):
            p = phlb[i, j]
            below = p[:-1] > edges[j, 0]
            strat = []
            for data in (strat_oh, strat_cl, strat_od):
                col = rebin(edges[j], data[j], p)
                strat.append(np.where(below, data[j, 0], col))
            cl_col = np.zeros(lm)
            if cl_trop is not None:
                te = np.concatenate([pl_trop[i, j], [0.0]])
                col = rebin(te, cl_trop[i, j], p)
                col = np.where(p[:-1] > te[0], cl_trop[i, j, 0], col)
                cl_col = (col * krate_cl[i, j]) * cl_scale
            strato = p[:-1] < trop[j]
            loss[i, j, :, 0] = np.where(strato, strat[0], oh_spiv[i, j])
            loss[i, j, :, 1] = np.where(strato, strat[1], strat[1] + cl_col)
            loss[i, j, :, 2] = strat[2] * o1d_scale
    if soil is not None:
        loss[..., 0, 3] = -(soil_rate * soil)
    return loss


def decay(loss, dt):
    total = ((loss[..., 0] + loss[..., 1]) + loss[..., 2]) + loss[..., 3]
    return libm.exp(-(total * dt))
