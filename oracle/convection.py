"""The 'v' operator: implicit convection plus Kzz diffusion per column (convection.F90 as
pydyn transcribes it, pydyn/kernels.py:350-457).

A column's tracer mass goes through the explicit inverse C of the (lmax, lmax) matrix conv1;
its slopes are scaled by diag(C).  The inverse is the Crout factorisation without pivoting and
substitution against the identity, each element accumulated in the reference's order.
"""

import numpy as np


def convection_matrix(m, entu, detu, entd, detd, cloud_top, cloud_lfs, dkg, dt, lmax):
    """conv1 (ncol, lmax, lmax) for columns m (ncol, >= lmax) ... (pydyn/kernels.py:350-396).
    Interface k (1-based, top of layer k) carries fractions fu/fd of each layer's tracer."""
    ncol = m.shape[0]
    fu = np.zeros((ncol, lmax + 1, lmax))
    fd = np.zeros((ncol, lmax + 1, lmax))
    amu = np.zeros((ncol, lmax + 1))
    amd = np.zeros((ncol, lmax + 1))
    with np.errstate(divide="ignore", invalid="ignore"):
        for k in range(1, lmax + 1):  # updraft, from the surface
            on = cloud_top >= k
            inflow = amu[:, k - 1] + entu[:, k - 1]
            a = inflow - detu[:, k - 1]
            rising = a > 0.0
            keep = np.where(rising, np.maximum(0.0, 1.0 - detu[:, k - 1] / inflow), 0.0)
            amu[:, k] = np.where(on & rising, a, 0.0)
            fu[:, k, : k - 1] = np.where(on[:, None], fu[:, k - 1, : k - 1] * keep[:, None], 0.0)
            fu[:, k, k - 1] = np.where(on, entu[:, k - 1] / m[:, k - 1] * keep, 0.0)
        for k in range(lmax, 1, -1):  # downdraft, from the level of free sinking
            on = cloud_lfs >= k
            inflow = amd[:, k] - entd[:, k - 1]
            a = inflow + detd[:, k - 1]
            sinking = a < 0.0
            keep = np.where(sinking, np.maximum(0.0, 1.0 + detd[:, k - 1] / inflow), 0.0)
            amd[:, k - 1] = np.where(on & sinking, a, 0.0)
            fd[:, k - 1, k:] = np.where(on[:, None], fd[:, k, k:] * keep[:, None], fd[:, k - 1, k:])
            fd[:, k - 1, k - 1] = np.where(on, -entd[:, k - 1] / m[:, k - 1] * keep, fd[:, k - 1, k - 1])
    f = fd
    for k in range(1, lmax):  # subsidence and diffusion through interface k
        f[:, k, :] = fu[:, k, :] + f[:, k, :]
        f[:, k, k] = f[:, k, k] - amu[:, k] / m[:, k]
        f[:, k, k - 1] = f[:, k, k - 1] - amd[:, k] / m[:, k - 1]
        f[:, k, k - 1] = f[:, k, k - 1] + dkg[:, k - 1] / m[:, k - 1]
        f[:, k, k] = f[:, k, k] - dkg[:, k - 1] / m[:, k]
    conv1 = np.empty((ncol, lmax, lmax))
    for k in range(lmax):
        conv1[:, k, :] = (-dt) * (f[:, k, :] - f[:, k + 1, :])
        conv1[:, k, k] = conv1[:, k, k] + 1.0
    return conv1


def crout_inverse(a):
    """The inverse of every (n, n) matrix of a by Crout LU without pivoting and substitution
    against the identity (convection.F90 fastminv, pydyn/kernels.py:399-420).  Every element is
    accumulated in increasing k, which is the reference's order.  The work is done on a
    columns-last copy (n, n, ncol) so that each statement runs over contiguous memory."""
    ncol, n, _ = a.shape
    lu = np.ascontiguousarray(np.moveaxis(a, 0, 2))
    for j in range(n):
        for k in range(j):
            lu[k + 1 :, j] = lu[k + 1 :, j] - lu[k + 1 :, k] * lu[k, j][None, :]
        lu[j + 1 :, j] = lu[j + 1 :, j] / lu[j, j][None, :]
    y = np.zeros((n, n, ncol))
    y[np.arange(n), np.arange(n)] = 1.0
    for i in range(n):
        for j in range(i):
            y[i] = y[i] - lu[i, j][None, :] * y[j]
    inv = np.empty_like(y)
    for i in range(n - 1, -1, -1):
        row = y[i]
        for j in range(i + 1, n):
            row = row - lu[i, j][None, :] * inv[j]
        inv[i] = row / lu[i, i][None, :]
    return np.moveaxis(inv, 2, 0)


def convec(m, rm, rxm, rym, rzm, entu, detu, entd, detd, cloud_top, cloud_lfs, dkg, dt, lmax, inverse="crout"):
    """Convection plus diffusion on every column of the region, in place (pydyn/kernels.py:
    423-457, on the cells a single region owns: all of them, pydyn/zoom/physics.py:494-513).
    m (im, jm, lm), tracers (im, jm, lm, nt), the meteo (im, jm, >= lmax), cloud levels (im, jm).
    inverse: 'crout' (fastminv) or 'lapack' (oracle/lapack.c: DGETRF2 + DGETRI, pydyn/lapack.py)."""
    im, jm = m.shape[:2]
    ncol = im * jm

    def col(q):
        return q.reshape(ncol, *q.shape[2:])

    conv1 = convection_matrix(
        col(m)[:, :lmax], col(entu), col(detu), col(entd), col(detd),
        col(cloud_top), col(cloud_lfs), col(dkg), dt, lmax,
    )  # fmt: skip
    if inverse == "lapack":
        from oracle import lapack

        c = lapack.inverse_rows(conv1)
    else:
        c = crout_inverse(conv1)
    diag = c[:, np.arange(lmax), np.arange(lmax)][:, :, None]
    for s in (rxm, rym, rzm):
        sc = col(s)
        sc[:, :lmax] = diag * sc[:, :lmax]
        s[...] = sc.reshape(s.shape)
    rmc = col(rm)
    old = rmc[:, :lmax].copy()
    acc = np.zeros_like(old)
    for k in range(lmax):
        acc = acc + c[:, :, k, None] * old[:, k, None, :]
    rmc[:, :lmax] = acc
    rm[...] = rmc.reshape(rm.shape)
