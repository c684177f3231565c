"""The 1-D pieces of the slopes scheme, on arrays whose axis 0 is the advected line.

Cells are axis 0 of length n, faces axis 0 of length n+1 (face k between cells k-1 and k).
Every expression keeps the reference's association order, since the results are compared bit
for bit (pydyn/ops.py:235-341, pydyn/kernels.py:64-92).
"""

import numpy as np

CFL_ONE = 1.0  # dims.F90 `one`, the threshold of every CFL test (pydyn/ops.py:44)


def seqsum(q):
    """Sum over axis 0 in index order, starting from the first term (ops.py:235-240)."""
    total = q[0].copy()
    for k in range(1, q.shape[0]):
        total = total + q[k]
    return total


def limit(s, rm):
    """The slope limiter clip(s, -|rm|, |rm|) (kernels.py:65-67)."""
    bound = np.abs(rm)
    return np.clip(s, -bound, bound)


def mass_after(m, f):
    """m + f(k-1) - f(k) for every cell, evaluated left to right (ops.py:257-259)."""
    return (m + f[:-1]) - f[1:]


def pick_upwind(f, lower, upper):
    """For each face flux f: `lower` where f >= 0 else `upper` (ops.py:262-264)."""
    return np.where(f >= 0.0, lower, upper)


def face_fluxes(flux, c, rm_u, s_u, t_u, w_u):
    """The slopes-scheme fluxes through faces: tracer mass f, its first moment pf, and the two
    cross slopes carried along (kernels.py:75-84).  Inputs are per face (upwind cell values)."""
    f = np.where(flux >= 0.0, c * (rm_u + (1.0 - c) * s_u), c * (rm_u - (1.0 + c) * s_u))
    pf = flux * ((c * c) * s_u - 3.0 * f)
    return f, pf, c * t_u, c * w_u


def slope_update(s, pf, flux, rm_new, f, m_old, m_new):
    """The along-flow slope after one sub-step, from the n+1 face values of pf, flux and f
    (kernels.py:87-92)."""
    lo, hi = flux[:-1], flux[1:]
    inner = ((pf[:-1] - pf[1:]) - (lo - hi) * s) + 3.0 * ((lo + hi) * rm_new - (f[:-1] + f[1:]) * m_old)
    return s + inner / m_new


def zero_ends(q):
    """q (n-1 interior faces, ...) with a zero face added at both ends (ops.py:281-300)."""
    out = np.zeros((q.shape[0] + 2, *q.shape[1:]))
    out[1:-1] = q
    return out


def periodic_ends(q):
    """q (n cells, ...) with the periodic halo: q(0) = q(n), q(n+1) = q(1) (ops.py:281-296)."""
    return np.concatenate([q[-1:], q, q[:1]], axis=0)


def shrink_progressive(f, n):
    """dynamu/dynamv's retry scaling f (n-1)/n of the current flux (ops.py:304-307)."""
    return f * (n - 1) / n


def shrink_fresh(f0, n):
    """dynamw's retry scaling of the original flux, f0 / n (ops.py:310-312)."""
    return f0 / float(n)


def nint(x):
    """Fortran nint: half away from zero (ops.py:78-80)."""
    r = np.floor(abs(x) + 0.5)
    return int(r) if x >= 0 else -int(r)
