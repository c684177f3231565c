"""dynam0 and the three slopes advection operators with their CFL sub-cycle searches.

Arrays are in the halo-free pydyn layout (C order): m (im, jm, lm), tracers (im, jm, lm, nt),
am (im+1, jm, lm), bm (im, jm+1, lm), cm (im, jm, lm+1).  Every operator works in place, as
the reference's do.  Lines are processed all at once (vectorised over rows / columns); each
line's arithmetic is exactly the one it gets alone, so the grouping does not change a bit.
"""

import numpy as np

from .lines import (
    CFL_ONE,
    face_fluxes,
    limit,
    mass_after,
    periodic_ends,
    pick_upwind,
    seqsum,
    shrink_fresh,
    shrink_progressive,
    slope_update,
    zero_ends,
)


class CFLError(RuntimeError):
    """A sub-cycle count exceeded its cap (pydyn/kernels.py:60-61)."""


# ------------------------------------------------------------------ dynam0
def dynam0(pu, pv, bt, dt):
    """am, bm, cm of a half step from the meteo fluxes pu (im+1, jm, lm), pv (im, jm+1, lm)
    (pydyn/kernels.py:102-125; the region-1 wrapper pydyn/zoom/massflow.py:446-461 adds the
    periodic face am(0) = am(im))."""
    im, jm, lm = pu.shape[0] - 1, pu.shape[1], pu.shape[2]
    conv = ((pu[:-1] - pu[1:]) + pv[:, :-1]) - pv[:, 1:]
    pit = seqsum(np.moveaxis(conv, 2, 0))
    cm = np.zeros((im, jm, lm + 1))
    sd = conv[..., lm - 1] - (bt[lm - 1] - bt[lm]) * pit
    cm[..., lm - 1] = (-dt) * sd
    for lev in range(lm - 2, 0, -1):
        sd = (sd + conv[..., lev]) - (bt[lev] - bt[lev + 1]) * pit
        cm[..., lev] = (-dt) * sd
    am = dt * pu
    am[0] = am[im]
    return am, dt * pv, cm


# ------------------------------------------------------------------ x: dynamu (periodic rows)
def periodic_search(m, a, nmax):
    """Each periodic row's smallest nloop whose nloop mass-only trial sub-steps keep every
    |Courant| < 1; the faces of a row that retries are scaled by (n-1)/n in place.  m (n, R),
    a (n+1, R).  Returns nloop (R,) or None once a row would need more than nmax
    (pydyn/kernels.py:169-199)."""
    rows = m.shape[1]
    nloop = np.ones(rows, dtype=np.int64)
    todo = np.arange(rows)
    n = 1
    while True:
        mx, f = m[:, todo].copy(), a[:, todo]
        passed = np.ones(todo.size, dtype=bool)
        for _ in range(n):
            ext = periodic_ends(mx)
            courant = np.abs(f / pick_upwind(f, ext[:-1], ext[1:]))
            passed &= np.all(courant < CFL_ONE, axis=0)
            mx = mass_after(mx, f)
        nloop[todo] = n
        todo = todo[~passed]
        if todo.size == 0:
            return nloop
        n += 1
        a[:, todo] = shrink_progressive(a[:, todo], n)
        if n > nmax:
            return None


def periodic_substep(m, rm, rxm, rym, rzm, a, limits):
    """One dynamu sub-step of periodic rows (pydyn/kernels.py:202-219): alpha = am * (1/m)."""
    af = a[..., None]
    m_new = mass_after(m, a)
    minv = periodic_ends(1.0 / m)[..., None]
    alpha = af * pick_upwind(af, minv[:-1], minv[1:])
    ups = []
    for q in (rm, rxm, rym, rzm):
        e = periodic_ends(q)
        ups.append(pick_upwind(af, e[:-1], e[1:]))
    f, pf, fy, fz = face_fluxes(af, alpha, *ups)
    for q in (f, pf, fy, fz):  # the seam: face 0 is face n
        q[0] = q[-1]
    rm_new = rm - (f[1:] - f[:-1])
    rxm_new = slope_update(rxm, pf, af, rm_new, f, m[..., None], m_new[..., None])
    if limits:
        rxm_new = limit(rxm_new, rm_new)
    return m_new, rm_new, rxm_new, rym - (fy[1:] - fy[:-1]), rzm - (fz[1:] - fz[:-1])


def advect_x(m, rm, rxm, rym, rzm, am, limits=True, max_nloop=50):
    """dynamu on every (j, l) row, each row with its own nloop (cap max_nloop - 1), the row's
    faces left scaled by the search and multiplied back by nloop (pydyn/kernels.py:129-166).
    Returns the per-row nloop (jm, lm)."""
    im, jm, lm = m.shape
    nt = rm.shape[3]
    rows = jm * lm
    mm = m.reshape(im, rows).copy()
    qs = [q.reshape(im, rows, nt).copy() for q in (rm, rxm, rym, rzm)]
    a = am.reshape(im + 1, rows).copy()
    if limits:
        qs[1] = limit(qs[1], qs[0])
    nloop = periodic_search(mm, a, max_nloop - 1)
    if nloop is None:
        raise CFLError("dynamu: nloop too high")
    for s in range(int(nloop.max())):
        sel = np.flatnonzero(nloop > s)
        parts = periodic_substep(mm[:, sel], *(q[:, sel] for q in qs), a[:, sel], limits)
        mm[:, sel] = parts[0]
        for q, p in zip(qs, parts[1:]):
            q[:, sel] = p
    a = a * nloop[None, :]
    m[...] = mm.reshape(m.shape)
    for dst, q in zip((rm, rxm, rym, rzm), qs):
        dst[...] = q.reshape(dst.shape)
    am[...] = a.reshape(am.shape)
    return nloop.reshape(jm, lm)


# ------------------------------------------------------------------ closed lines (y, z)
def closed_courant_ok(mx, f):
    """Every interior face's |Courant| < 1 for the whole array (axis 0 the line)."""
    fi = f[1:-1]
    return bool(np.all(np.abs(fi / pick_upwind(fi, mx[:-1], mx[1:])) < CFL_ONE))


def closed_search(m, f, nmax, fresh):
    """ops.subcycles on closed lines that share one nloop (pydyn/ops.py:315-341): the smallest
    n whose trial passes on every line.  progressive (dynamv): f scaled in place, cumulatively;
    fresh (dynamw): the original divided by n, f left alone.  Returns (n, flux used) or None."""
    f0, n = f, 1
    while True:
        mx, ok = m.copy(), True
        for _ in range(n):
            if not closed_courant_ok(mx, f):
                ok = False
                break
            mx = mass_after(mx, f)
        if ok:
            return n, f
        n += 1
        if fresh:
            f = shrink_fresh(f0, n)
        else:
            f[...] = shrink_progressive(f, n)
        if n > nmax:
            return None


def _y_layer(m, rm, rxm, rym, rzm, b, limits, nmax):
    """dynamv on one layer, the y axis first: m (jm, im), tracers (jm, im, nt), b (jm+1, im),
    all in place (pydyn/kernels.py:246-291).  Returns nloop."""
    if limits:
        rym[...] = limit(rym, rm)
    found = closed_search(m, b, nmax, fresh=False)
    if found is None:
        raise CFLError("dynamv: nloop too high")
    nloop = found[0]
    bf = b[1:-1, :, None]
    for _ in range(nloop):
        m_new = mass_after(m, b)
        beta = bf / pick_upwind(bf, m[:-1, :, None], m[1:, :, None])
        ups = [pick_upwind(bf, q[:-1], q[1:]) for q in (rm, rym, rxm, rzm)]
        f, pf, fx, fz = face_fluxes(bf, beta, *ups)
        # flow out of a pole cell carries no slope (dynamv's faces 2 and jm)
        for face, cell, leaving in ((0, 0, bf[0] >= 0.0), (-1, -1, bf[-1] < 0.0)):
            f[face] = np.where(leaving, beta[face] * rm[cell], f[face])
            pf[face] = np.where(leaving, (-3.0 * bf[face]) * f[face], pf[face])
            fx[face] = np.where(leaving, 0.0, fx[face])
            fz[face] = np.where(leaving, beta[face] * rzm[cell], fz[face])
        F, PF, FX, FZ = (zero_ends(q) for q in (f, pf, fx, fz))
        keep_y, keep_x = rym[[0, -1]].copy(), rxm[[0, -1]].copy()
        inv_mnew = (1.0 / m_new)[:, :, None]
        flow_conv = (b[:-1] - b[1:])[:, :, None]
        flow_mean = (3.0 * (b[:-1] + b[1:]))[:, :, None]
        mass_x3 = (3.0 * m)[:, :, None]
        rm[...] = rm - (F[1:] - F[:-1])
        rym[...] = rym + ((((PF[:-1] - PF[1:]) - flow_conv * rym) + flow_mean * rm) - mass_x3 * (F[:-1] + F[1:])) * inv_mnew
        if limits:
            rym[...] = limit(rym, rm)
        rxm[...] = rxm - (FX[1:] - FX[:-1])
        rzm[...] = rzm - (FZ[1:] - FZ[:-1])
        rym[[0, -1]], rxm[[0, -1]] = keep_y, keep_x
        m[...] = m_new
    b[...] = b * nloop
    return nloop


def advect_y(m, rm, rxm, rym, rzm, bm, limits=True, max_nloop=10):
    """dynamv: meridional advection with polar caps, one nloop per layer (cap max_nloop - 1)
    (pydyn/kernels.py:223-243).  In place; returns the per-layer nloop (lm,)."""
    jm, lm = m.shape[1], m.shape[2]
    if np.any(bm[:, 0] != 0.0) or np.any(bm[:, jm] != 0.0):
        raise ValueError("advect_y: non-zero flux through the polar outer faces")
    out = np.zeros(lm, dtype=np.int64)
    for k in range(lm):
        views = [np.swapaxes(q[:, :, k], 0, 1) for q in (m, rm, rxm, rym, rzm, bm)]
        out[k] = _y_layer(*views, limits, max_nloop - 1)
    return out


def z_substeps(m, rm, rxm, rym, rzm, cm, n_advz, limits=True):
    """dynamw's n_advz sub-steps with cm / n_advz; cm is not changed (the reference divides a
    copy and restores it exactly) (pydyn/kernels.py:324-346)."""
    c = np.moveaxis(cm, 2, 0) / float(n_advz)
    m_, rm_, rxm_, rym_, rzm_ = (np.moveaxis(q, 2, 0) for q in (m, rm, rxm, rym, rzm))
    cf = c[1:-1, ..., None]
    for _ in range(n_advz):
        m_new = mass_after(m_, c)
        gamma = cf / pick_upwind(cf, m_[:-1, ..., None], m_[1:, ..., None])
        ups = [pick_upwind(cf, q[:-1], q[1:]) for q in (rm_, rzm_, rxm_, rym_)]
        f, pf, fx, fy = face_fluxes(cf, gamma, *ups)
        F, PF, FX, FY = (zero_ends(q) for q in (f, pf, fx, fy))
        rm_[...] = mass_after(rm_, F)
        rzm_[...] = slope_update(rzm_, PF, c[..., None], rm_, F, m_[..., None], m_new[..., None])
        if limits:
            rzm_[...] = limit(rzm_, rm_)
        rxm_[...] = mass_after(rxm_, FX)
        rym_[...] = mass_after(rym_, FY)
        m_[...] = m_new


def advect_z(m, rm, rxm, rym, rzm, cm, limits=True, iter_limit=15):
    """dynamw: vertical advection, one n_advz for the whole region (cap iter_limit), cm / n
    taken fresh (pydyn/kernels.py:295-321).  In place; returns n_advz."""
    lm = m.shape[2]
    if np.any(cm[..., 0] != 0.0) or np.any(cm[..., lm] != 0.0):
        raise ValueError("dynamw: non-zero halo interfaces in cm")
    if limits:
        rzm[...] = limit(rzm, rm)
    found = closed_search(np.moveaxis(m, 2, 0), np.moveaxis(cm, 2, 0), iter_limit, fresh=True)
    if found is None:
        raise CFLError("dynamw: too many CFL violations!")
    n_advz = found[0]
    z_substeps(m, rm, rxm, rym, rzm, cm, n_advz, limits)
    return n_advz
