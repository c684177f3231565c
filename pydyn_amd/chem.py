"""The host side of the CH4 loss-rate construction (pydyn/chem.py:106-119): the zonal columns'
edges and the tropopause pressure, which the reference also computes on the host before its
device kernel (pydyn/gpu/kernels.py:598-599); the rates themselves are tm5_loss_rates."""

import math

import numpy as np

TROP_P0, TROP_DP = 30000.0, 21500.0  # tropopause_pressure = 30000 - 21500 cos^2(lat), Pa


def strat_edges(press):
    """Edges (…, n+1) in Pa of the zonal columns given at pressures `press` (…, n) hPa: half a
    layer below the first, the midpoints, 0 on top."""
    press = np.asarray(press, dtype=np.float64)
    n = press.shape[-1]
    edges = np.empty((*press.shape[:-1], n + 1))
    edges[..., 0] = press[..., 0] + 0.5 * (press[..., 0] - press[..., 1])
    edges[..., 1:n] = (press[..., :-1] + press[..., 1:]) / 2.0
    edges[..., n] = 0.0
    return edges * 100.0


def tropopause(lat):
    """Pressure (Pa) of the tropopause at latitudes `lat` (radians), cos from the C library (the
    reference's libm route for cos)."""
    lat = np.asarray(lat, dtype=np.float64)
    c = np.array([math.cos(v) for v in lat.ravel()]).reshape(lat.shape)
    return TROP_P0 - TROP_DP * (c * c)
