"""Tangent-linear (TLM) and adjoint (ADM) models of the transport step's tracer map -- TEST
INFRASTRUCTURE, the NumPy reference the device adjoint is checked against.

Given the air-mass trajectory (which the tracers do not influence), every operator maps the
tracers (rm, rxm, rym, rzm) linearly, except for the slope limiter clip(s, -|rm|, |rm|), which
is piecewise linear: its derivative is d s where it passes s, and +-sign(rm) d rm where it
clips.  The TLM propagates perturbations through each operator of a recorded forward window
(oracle.driver.Run), taking the limiter's branch from the recomputed forward state; the ADM
applies the transposes in reverse.  The controls are the initial tracers and the emission field
(the injection is affine in it); the loss is the sum of squared misfits of column-average
tracer-1 mixing ratios at sample columns.

Each operator follows the forward restatement it linearises (oracle/advection.py, convection.py,
physics.py; pydyn/kernels.py:129-457, pydyn/emis.py:52-54, pydyn/chem.py:168-178).
"""

import hashlib

import numpy as np

from . import physics
from .advection import closed_search, periodic_search
from .convection import convection_matrix, crout_inverse
from .driver import SEC_MONTH, XMCH4, Run
from .lines import face_fluxes, limit, mass_after, periodic_ends, pick_upwind, slope_update

FIELDS = ("rm", "rxm", "rym", "rzm")


# ------------------------------------------------------------------ the limiter
def lim_branch(s, r):
    """The branch np.clip(s, -|r|, |r|) takes (max by `>` then min by `<`): 0 s, 1 -|r|, 2 |r|."""
    hi = np.abs(r)
    lo = -hi
    low = ~(s > lo)
    t = np.where(low, lo, s)
    high = ~(t < hi)
    return np.where(high, 2, np.where(low, 1, 0))


def lim_tl(code, ds, dr, r):
    sg = np.sign(r)
    return np.where(code == 0, ds, np.where(code == 1, -sg * dr, sg * dr))


def lim_ad(code, lam, r):
    """(lambda of s, lambda of r) from lambda of the clipped value."""
    sg = np.sign(r)
    return np.where(code == 0, lam, 0.0), np.where(code == 1, -sg * lam, np.where(code == 2, sg * lam, 0.0))


# ------------------------------------------------------------------ x: periodic rows
def _x_coeffs(m, a):
    """Upwind cells and Courant numbers of a periodic sub-step: m (n, R), a (n+1, R)."""
    n = m.shape[0]
    k = np.arange(n + 1)[:, None]
    pos = a >= 0.0
    up = np.where(pos, (k - 1) % n, k % n)  # (n+1, R) the upwind cell of each face
    minv = periodic_ends(1.0 / m)
    alpha = a * pick_upwind(a, minv[:-1], minv[1:])
    return pos, up, alpha


def _gather(q, up):
    """q (n, R, nt) at the cells up (F, R) -> (F, R, nt)."""
    return np.take_along_axis(q, up[..., None].repeat(q.shape[2], axis=2), axis=0)


def _scatter(q, up, v):
    """q[up[k], r] += v[k, r] for every face k (accumulated, repeats allowed)."""
    n, rows, nt = q.shape
    cols = np.broadcast_to(np.arange(rows)[None, :], up.shape)
    for t in range(nt):
        np.add.at(q[..., t], (up, cols), v[..., t])


def x_forward_tl(m, a, Q, D, limits):
    """One periodic sub-step of rows: forward state Q -> new state, and its TLM D -> D'.
    Returns (m_new, Q_new, D_new, record) with what the ADM needs."""
    pos, up, alpha = _x_coeffs(m, a)
    af, al = a[..., None], alpha[..., None]
    m_new = mass_after(m, a)
    ups = [_gather(q, up) for q in Q]
    f, pf, ft, fw = face_fluxes(af, al, *ups)
    for q in (f, pf, ft, fw):
        q[0] = q[-1]
    rm_new = Q[0] - (f[1:] - f[:-1])
    s_raw = slope_update(Q[1], pf, af, rm_new, f, m[..., None], m_new[..., None])
    code = lim_branch(s_raw, rm_new) if limits else np.zeros(s_raw.shape, dtype=int)
    s_new = limit(s_raw, rm_new) if limits else s_raw
    Qn = [rm_new, s_new, Q[2] - (ft[1:] - ft[:-1]), Q[3] - (fw[1:] - fw[:-1])]
    # the same operations on the perturbations
    dups = [_gather(d, up) for d in D]
    g = np.where(af >= 0.0, al * (1.0 - al), -(al * (1.0 + al)))
    df = al * dups[0] + g * dups[1]
    dpf = af * ((al * al) * dups[1] - 3.0 * df)
    dft, dfw = al * dups[2], al * dups[3]
    for q in (df, dpf, dft, dfw):
        q[0] = q[-1]
    drm = D[0] - (df[1:] - df[:-1])
    lo, hi = af[:-1], af[1:]
    ds_raw = D[1] + (((dpf[:-1] - dpf[1:]) - (lo - hi) * D[1]) + 3.0 * ((lo + hi) * drm - (df[:-1] + df[1:]) * m[..., None])) / m_new[..., None]
    ds = lim_tl(code, ds_raw, drm, rm_new) if limits else ds_raw
    Dn = [drm, ds, D[2] - (dft[1:] - dft[:-1]), D[3] - (dfw[1:] - dfw[:-1])]
    rec = (m, m_new, a, up, alpha, code, rm_new)
    return m_new, Qn, Dn, rec


def x_ad(rec, L, limits):
    """The transpose of x_forward_tl's perturbation map: lambdas after -> lambdas before."""
    m, m_new, a, up, alpha, code, rm_new = rec
    af, al = a[..., None], alpha[..., None]
    n = m.shape[0]
    lrm, ls_new, lt_new, lw_new = (x.copy() for x in L)
    if limits:
        ls_raw, lr = lim_ad(code, ls_new, rm_new)
        lrm = lrm + lr
    else:
        ls_raw = ls_new
    lo, hi = af[:-1], af[1:]
    qv = ls_raw / m_new[..., None]
    ls = ls_raw - (lo - hi) * qv
    shape = (n + 1, *lrm.shape[1:])
    lpf, lf, lft, lfw = (np.zeros(shape) for _ in range(4))
    lpf[:-1] += qv
    lpf[1:] -= qv
    lrm = lrm + 3.0 * (lo + hi) * qv
    lf[:-1] += -3.0 * m[..., None] * qv
    lf[1:] += -3.0 * m[..., None] * qv
    lf[1:] -= lrm
    lf[:-1] += lrm
    lft[1:] -= lt_new
    lft[:-1] += lt_new
    lfw[1:] -= lw_new
    lfw[:-1] += lw_new
    for q in (lf, lpf, lft, lfw):  # the seam: face 0 carries face n's values
        q[-1] += q[0]
        q[0] = 0.0
    ls_u = af * (al * al) * lpf
    lf = lf + (-3.0 * af) * lpf
    g = np.where(af >= 0.0, al * (1.0 - al), -(al * (1.0 + al)))
    lrm_u = al * lf
    ls_u = ls_u + g * lf
    lt_u, lw_u = al * lft, al * lfw
    out = [lrm, ls, lt_new.copy(), lw_new.copy()]
    for q, v in zip(out, (lrm_u, ls_u, lt_u, lw_u)):
        _scatter(q, up[1:], v[1:])
    return out


def x_op(m, am, Q, D, limits, mode, max_nloop=50):
    """The x operator's TLM (mode 'tl': D perturbations in, out) or ADM (mode 'ad': D lambdas
    after, out before), the forward inputs m, am, Q = (rm, rxm, rym, rzm) recorded before it."""
    im, jm, lm = m.shape
    nt = Q[0].shape[3]
    R = jm * lm
    mm = m.reshape(im, R).copy()
    a = am.reshape(im + 1, R).copy()
    q = [x.reshape(im, R, nt).copy() for x in Q]
    d = [x.reshape(im, R, nt).copy() for x in D]
    code0 = lim_branch(q[1], q[0]) if limits else None
    r0 = q[0].copy()
    if limits:
        q[1] = limit(q[1], q[0])
    nloop = periodic_search(mm.copy(), a, max_nloop - 1)
    recs = []
    if mode == "tl" and limits:
        d[1] = lim_tl(code0, d[1], d[0], r0)
    for s in range(int(nloop.max())):
        sel = np.flatnonzero(nloop > s)
        parts = [x[:, sel] for x in q]
        dparts = [x[:, sel] for x in d] if mode == "tl" else [np.zeros_like(x[:, sel]) for x in q]
        m_new, qn, dn, rec = x_forward_tl(mm[:, sel], a[:, sel], parts, dparts, limits)
        mm[:, sel] = m_new
        for x, p in zip(q, qn):
            x[:, sel] = p
        if mode == "tl":
            for x, p in zip(d, dn):
                x[:, sel] = p
        recs.append((sel, rec))
    if mode == "ad":
        for sel, rec in reversed(recs):
            lam = x_ad(rec, [x[:, sel] for x in d], limits)
            for x, p in zip(d, lam):
                x[:, sel] = p
        if limits:
            ls, lr = lim_ad(code0, d[1], r0)
            d[1] = ls
            d[0] = d[0] + lr
    return [x.reshape(Q[0].shape) for x in d]


# ------------------------------------------------------------------ y: closed lines with polar caps
def y_forward_tl(m, b, Q, D, limits):
    """One dynamv sub-step of lines (jm cells, axis 0) with the polar rules, and its TLM."""
    bf = b[1:-1, ..., None]
    mlo, mhi = m[:-1, ..., None], m[1:, ..., None]
    beta = bf / pick_upwind(bf, mlo, mhi)
    m_new = mass_after(m, b)

    def fluxes(X, tl):
        rm, s, t, w = X
        ups = [pick_upwind(bf, x[:-1], x[1:]) for x in (rm, s, t, w)]
        if not tl:
            f, pf, ft, fw = face_fluxes(bf, beta, *ups)
        else:
            g = np.where(bf >= 0.0, beta * (1.0 - beta), -(beta * (1.0 + beta)))
            f = beta * ups[0] + g * ups[1]
            pf = bf * ((beta * beta) * ups[1] - 3.0 * f)
            ft, fw = beta * ups[2], beta * ups[3]
        for face, cell, leaving in ((0, 0, bf[0] >= 0.0), (-1, -1, bf[-1] < 0.0)):
            f[face] = np.where(leaving, beta[face] * rm[cell], f[face])
            pf[face] = np.where(leaving, (-3.0 * bf[face]) * f[face], pf[face])
            ft[face] = np.where(leaving, 0.0, ft[face])
            fw[face] = np.where(leaving, beta[face] * w[cell], fw[face])
        return [np.concatenate([np.zeros_like(x[:1]), x, np.zeros_like(x[:1])]) for x in (f, pf, ft, fw)]

    def update(X, F, tl, code=None, rm_fwd=None):
        rm, s, t, w = X
        Fm, PF, FT, FW = F
        inv = (1.0 / m_new)[..., None]
        conv = (b[:-1] - b[1:])[..., None]
        mean = (3.0 * (b[:-1] + b[1:]))[..., None]
        mx3 = (3.0 * m)[..., None]
        rm_new = rm - (Fm[1:] - Fm[:-1])
        s_raw = s + ((((PF[:-1] - PF[1:]) - conv * s) + mean * rm_new) - mx3 * (Fm[:-1] + Fm[1:])) * inv
        return rm_new, s_raw, t - (FT[1:] - FT[:-1]), w - (FW[1:] - FW[:-1])

    F = fluxes(Q, False)
    rm_new, s_raw, t_new, w_new = update(Q, F, False)
    code = lim_branch(s_raw, rm_new) if limits else np.zeros(s_raw.shape, dtype=int)
    s_new = limit(s_raw, rm_new) if limits else s_raw.copy()
    s_new[[0, -1]], t_new[[0, -1]] = Q[1][[0, -1]], Q[2][[0, -1]]
    dF = fluxes(D, True)
    drm, ds_raw, dt_new, dw_new = update(D, dF, True)
    ds = lim_tl(code, ds_raw, drm, rm_new) if limits else ds_raw.copy()
    ds[[0, -1]], dt_new[[0, -1]] = D[1][[0, -1]], D[2][[0, -1]]
    rec = (m, m_new, b, beta, code, rm_new)
    return m_new, [rm_new, s_new, t_new, w_new], [drm, ds, dt_new, dw_new], rec


def y_ad(rec, L, limits):
    m, m_new, b, beta, code, rm_new = rec
    bf = b[1:-1, ..., None]
    lrm, ls_new, lt_new, lw_new = (x.copy() for x in L)
    # the pole cells keep their rym, rxm: their new values are the old ones
    keep_s, keep_t = ls_new[[0, -1]].copy(), lt_new[[0, -1]].copy()
    ls_new[[0, -1]] = 0.0
    lt_new[[0, -1]] = 0.0
    if limits:
        ls_raw, lr = lim_ad(code, ls_new, rm_new)
        lrm = lrm + lr
    else:
        ls_raw = ls_new
    inv = (1.0 / m_new)[..., None]
    conv = (b[:-1] - b[1:])[..., None]
    mean = (3.0 * (b[:-1] + b[1:]))[..., None]
    mx3 = (3.0 * m)[..., None]
    qv = ls_raw * inv
    ls = ls_raw - conv * qv
    nf = b.shape[0]
    shape = (nf, *lrm.shape[1:])
    LF, LPF, LFT, LFW = (np.zeros(shape) for _ in range(4))
    LPF[:-1] += qv
    LPF[1:] -= qv
    lrm = lrm + mean * qv
    LF[:-1] += -mx3 * qv
    LF[1:] += -mx3 * qv
    LF[1:] -= lrm
    LF[:-1] += lrm
    LFT[1:] -= lt_new
    LFT[:-1] += lt_new
    LFW[1:] -= lw_new
    LFW[:-1] += lw_new
    lf, lpf, lft, lfw = (x[1:-1] for x in (LF, LPF, LFT, LFW))  # the closed faces carry nothing
    lrm_out, ls_out, lt_out, lw_out = lrm.copy(), ls, lt_new.copy(), lw_new.copy()
    lt_out[[0, -1]] += keep_t
    ls_out[[0, -1]] += keep_s
    # polar faces: where the flow leaves a pole cell, f = beta rm(cell), pf = -3 bf f, ft = 0, fw = beta w(cell)
    leave = [(0, 0, bf[0] >= 0.0), (-1, -1, bf[-1] < 0.0)]
    special = np.zeros(lf.shape, dtype=bool)
    for face, cell, leaving in leave:
        lfv = lf[face] + (-3.0 * bf[face]) * lpf[face]
        lrm_out[cell] += np.where(leaving, beta[face] * lfv, 0.0)
        lw_out[cell] += np.where(leaving, beta[face] * lfw[face], 0.0)
        special[face] = np.broadcast_to(leaving, special[face].shape)
    reg = ~special
    lpf_r, lf_r = np.where(reg, lpf, 0.0), np.where(reg, lf, 0.0)
    lft_r, lfw_r = np.where(reg, lft, 0.0), np.where(reg, lfw, 0.0)
    lf_r = lf_r + (-3.0 * bf) * lpf_r
    g = np.where(bf >= 0.0, beta * (1.0 - beta), -(beta * (1.0 + beta)))
    contrib = [beta * lf_r, bf * (beta * beta) * lpf_r + g * lf_r, beta * lft_r, beta * lfw_r]
    pos = bf >= 0.0
    for out, v in zip((lrm_out, ls_out, lt_out, lw_out), contrib):
        out[:-1] += np.where(pos, v, 0.0)
        out[1:] += np.where(pos, 0.0, v)
    return [lrm_out, ls_out, lt_out, lw_out]


def y_op(m, bm, Q, D, limits, mode, max_nloop=10):
    """The y operator's TLM / ADM per layer (the layer's line axis first)."""
    lm = m.shape[2]
    out = [np.empty_like(x) for x in D]
    for k in range(lm):
        mm = np.swapaxes(m[:, :, k], 0, 1).copy()
        b = np.swapaxes(bm[:, :, k], 0, 1).copy()
        q = [np.swapaxes(x[:, :, k], 0, 1).copy() for x in Q]  # (jm, im, nt), s = rym, t = rxm
        q = [q[0], q[2], q[1], q[3]]
        d = [np.swapaxes(x[:, :, k], 0, 1).copy() for x in D]
        d = [d[0], d[2], d[1], d[3]]
        code0 = lim_branch(q[1], q[0]) if limits else None
        r0 = q[0].copy()
        if limits:
            q[1] = limit(q[1], q[0])
            if mode == "tl":
                d[1] = lim_tl(code0, d[1], d[0], r0)
        found = closed_search(mm, b, max_nloop - 1, fresh=False)
        recs = []
        for _ in range(found[0]):
            m_new, q, dn, rec = y_forward_tl(mm, b, q, d if mode == "tl" else [np.zeros_like(x) for x in q], limits)
            if mode == "tl":
                d = dn
            recs.append(rec)
            mm = m_new
        if mode == "ad":
            for rec in reversed(recs):
                d = y_ad(rec, d, limits)
            if limits:
                ls, lr = lim_ad(code0, d[1], r0)
                d[1], d[0] = ls, d[0] + lr
        d = [d[0], d[2], d[1], d[3]]
        for o, x in zip(out, d):
            o[:, :, k] = np.swapaxes(x, 0, 1)
    return out


# ------------------------------------------------------------------ z: closed columns, n_advz
def z_forward_tl(m, c, Q, D, limits):
    cf = c[1:-1, ..., None]
    gamma = cf / pick_upwind(cf, m[:-1, ..., None], m[1:, ..., None])
    m_new = mass_after(m, c)

    def fluxes(X, tl):
        rm, s, t, w = X
        ups = [pick_upwind(cf, x[:-1], x[1:]) for x in (rm, s, t, w)]
        if not tl:
            out = face_fluxes(cf, gamma, *ups)
        else:
            g = np.where(cf >= 0.0, gamma * (1.0 - gamma), -(gamma * (1.0 + gamma)))
            f = gamma * ups[0] + g * ups[1]
            out = (f, cf * ((gamma * gamma) * ups[1] - 3.0 * f), gamma * ups[2], gamma * ups[3])
        return [np.concatenate([np.zeros_like(x[:1]), x, np.zeros_like(x[:1])]) for x in out]

    def update(X, F):
        rm, s, t, w = X
        rm_new = mass_after(rm, F[0])
        s_raw = slope_update(s, F[1], c[..., None], rm_new, F[0], m[..., None], m_new[..., None])
        return rm_new, s_raw, mass_after(t, F[2]), mass_after(w, F[3])

    rm_new, s_raw, t_new, w_new = update(Q, fluxes(Q, False))
    code = lim_branch(s_raw, rm_new) if limits else np.zeros(s_raw.shape, dtype=int)
    s_new = limit(s_raw, rm_new) if limits else s_raw
    drm, ds_raw, dt_new, dw_new = update(D, fluxes(D, True))
    ds = lim_tl(code, ds_raw, drm, rm_new) if limits else ds_raw
    rec = (m, m_new, c, gamma, code, rm_new)
    return m_new, [rm_new, s_new, t_new, w_new], [drm, ds, dt_new, dw_new], rec


def z_ad(rec, L, limits):
    m, m_new, c, gamma, code, rm_new = rec
    cf = c[1:-1, ..., None]
    lrm, ls_new, lt_new, lw_new = (x.copy() for x in L)
    if limits:
        ls_raw, lr = lim_ad(code, ls_new, rm_new)
        lrm = lrm + lr
    else:
        ls_raw = ls_new
    lo, hi = c[:-1, ..., None], c[1:, ..., None]
    qv = ls_raw / m_new[..., None]
    ls = ls_raw - (lo - hi) * qv
    nf = c.shape[0]
    shape = (nf, *lrm.shape[1:])
    LF, LPF, LFT, LFW = (np.zeros(shape) for _ in range(4))
    LPF[:-1] += qv
    LPF[1:] -= qv
    lrm = lrm + 3.0 * (lo + hi) * qv
    LF[:-1] += -3.0 * m[..., None] * qv
    LF[1:] += -3.0 * m[..., None] * qv
    # rm_new = rm + F_lo - F_hi; t, w likewise
    LF[:-1] += lrm
    LF[1:] -= lrm
    LFT[:-1] += lt_new
    LFT[1:] -= lt_new
    LFW[:-1] += lw_new
    LFW[1:] -= lw_new
    lf, lpf, lft, lfw = (x[1:-1] for x in (LF, LPF, LFT, LFW))
    lf = lf + (-3.0 * cf) * lpf
    g = np.where(cf >= 0.0, gamma * (1.0 - gamma), -(gamma * (1.0 + gamma)))
    contrib = [gamma * lf, cf * (gamma * gamma) * lpf + g * lf, gamma * lft, gamma * lfw]
    out = [lrm, ls, lt_new.copy(), lw_new.copy()]
    pos = cf >= 0.0
    for o, v in zip(out, contrib):
        o[:-1] += np.where(pos, v, 0.0)
        o[1:] += np.where(pos, 0.0, v)
    return out


def z_op(m, cm, Q, D, limits, mode, iter_limit=15):
    """The z operator's TLM / ADM (levels first; s = rzm, t = rxm, w = rym)."""
    mz = np.moveaxis(m, 2, 0).copy()
    cz = np.moveaxis(cm, 2, 0)
    q = [np.moveaxis(x, 2, 0).copy() for x in Q]
    d = [np.moveaxis(x, 2, 0).copy() for x in D]
    q, d = [q[0], q[3], q[1], q[2]], [d[0], d[3], d[1], d[2]]
    code0 = lim_branch(q[1], q[0]) if limits else None
    r0 = q[0].copy()
    if limits:
        q[1] = limit(q[1], q[0])
        if mode == "tl":
            d[1] = lim_tl(code0, d[1], d[0], r0)
    n = closed_search(mz, cz, iter_limit, fresh=True)[0]
    c = cz / float(n)
    recs = []
    for _ in range(n):
        mz_new, q, dn, rec = z_forward_tl(mz, c, q, d if mode == "tl" else [np.zeros_like(x) for x in q], limits)
        if mode == "tl":
            d = dn
        recs.append(rec)
        mz = mz_new
    if mode == "ad":
        for rec in reversed(recs):
            d = z_ad(rec, d, limits)
        if limits:
            ls, lr = lim_ad(code0, d[1], r0)
            d[1], d[0] = ls, d[0] + lr
    d = [d[0], d[2], d[3], d[1]]
    return [np.moveaxis(x, 0, 2) for x in d]


# ------------------------------------------------------------------ column physics
_INVERSES = {}  # the last convection inverses, by the bytes of m and dt (a step's two v share them)


def v_op(m, met, dkg, dt, lmax, D, mode):
    """convec: rm <- C rm over the lowest lmax levels, slopes times diag(C); the ADM uses C^T."""
    im, jm = m.shape[:2]
    ncol = im * jm

    def col(q):
        return q.reshape(ncol, *q.shape[2:])

    key = (hashlib.sha1(np.ascontiguousarray(m).tobytes()).hexdigest(), float(dt), lmax, id(met))
    C = _INVERSES.get(key)
    if C is None:
        conv1 = convection_matrix(col(m)[:, :lmax], col(met.entu), col(met.detu), col(met.entd), col(met.detd),
                                  col(met.cloud_top), col(met.cloud_lfs), col(dkg), dt, lmax)  # fmt: skip
        C = crout_inverse(conv1)  # (ncol, lmax, lmax)
        _INVERSES.clear()
        _INVERSES[key] = C
    out = [x.copy() for x in D]
    diag = C[:, np.arange(lmax), np.arange(lmax)][:, :, None]
    for k in (1, 2, 3):
        s = col(out[k])
        s[:, :lmax] = diag * s[:, :lmax]
    r = col(out[0])
    M = C if mode == "tl" else np.swapaxes(C, 1, 2)
    r[:, :lmax] = np.einsum("ckl,clt->ckt", M, r[:, :lmax])
    return out


# ------------------------------------------------------------------ the window
class Window:
    """A forward window of the single-region run (oracle.driver.Run) with every tracer
    operator's inputs recorded: [(op, ndyn, m, flux, tracers)]."""

    def __init__(self, met, intervals, limits=True, emis=None, init=None, on_op=None, progress=None):
        """on_op: called with every operator's record as the forward reaches it, instead of the
        record being kept in ops (a tangent run alongside the forward, tangent()); progress:
        called with each interval's index once it has run."""
        self.met, self.limits = met, limits
        self.run = Run(met, limits)
        if init is not None:
            for k, v in zip(FIELDS, init):
                self.run.st[k] = v.copy()
        self.emis = met.emis if emis is None else emis
        self.ops = []
        run, window = self.run, self

        def operator(op, ndyn):
            st = run.st
            flux = {"x": "am", "y": "bm", "z": "cm"}.get(op)
            rec = (op, ndyn, st["m"].copy(), None if flux is None else st[flux].copy(), [st[k].copy() for k in FIELDS])
            if on_op is None:
                window.ops.append(rec)
            else:
                on_op(rec)
            if op == "s":
                factor = physics.month2dt(ndyn, SEC_MONTH, 1)
                x = physics.inject(window.emis, factor, XMCH4, XMCH4, 1.0)
                physics.add_2d(st["rm"], st["rzm"], x, 1, 1)
            else:
                Run.operator(run, op, ndyn)

        run.operator = operator
        for k in range(intervals):
            run.interval()
            if progress is not None:
                progress(k)
        self.final = [run.st[k].copy() for k in FIELDS]
        self.m_final = run.st["m"].copy()

    @classmethod
    def tangent(cls, met, intervals, d_init, d_emis, limits=True, emis=None, init=None, progress=None):
        """The TLM run alongside the forward, operator by operator, nothing recorded: (the
        perturbations of the final tracers, the window, whose final and m_final are the
        forward's)."""
        w = cls.__new__(cls)
        D = [x.copy() for x in d_init]

        def step(rec):  # _op needs met, limits, emis and run: the constructor sets them first
            nonlocal D
            op, ndyn, m, flux, Q = rec
            D = w._op(op, ndyn, m, flux, Q, D, "tl", d_emis)

        cls.__init__(w, met, intervals, limits=limits, emis=emis, init=init, on_op=step, progress=progress)
        return D, w

    def tlm(self, d_init, d_emis):
        """Perturbations of the final tracers from those of the initial tracers and the emission."""
        D = [x.copy() for x in d_init]
        met, g = self.met, self.met.g
        for op, ndyn, m, flux, Q in self.ops:
            D = self._op(op, ndyn, m, flux, Q, D, "tl", d_emis)
        return D

    def adm(self, L_final):
        """(lambdas of the initial tracers, lambda of the emission) from those of the final ones."""
        L = [x.copy() for x in L_final]
        lam_emis = np.zeros_like(self.emis)
        for op, ndyn, m, flux, Q in reversed(self.ops):
            L = self._op(op, ndyn, m, flux, Q, L, "ad", lam_emis)
        return L, lam_emis

    def _op(self, op, ndyn, m, flux, Q, D, mode, emis_slot):
        met, g = self.met, self.met.g
        if op == "x":
            return x_op(m, flux, Q, D, self.limits, mode)
        if op == "y":
            return y_op(m, flux, Q, D, self.limits, mode)
        if op == "z":
            return z_op(m, flux, Q, D, self.limits, mode)
        if op == "v":
            return v_op(m, met, self.run.dkg, ndyn / 2.0, g.lmax, D, mode)
        if op == "c":
            factor = physics.decay(met.loss, ndyn / 2.0)
            return [x * factor[..., None] for x in D]
        if op == "s":
            k = ((physics.month2dt(ndyn, SEC_MONTH, 1)) * (XMCH4 / XMCH4)) * 1.0
            D = [x.copy() for x in D]
            if mode == "tl":
                D[0][:, :, 0, 0] += k * emis_slot
                D[3][:, :, 0, 0] -= k * emis_slot
            else:
                emis_slot += k * (D[0][:, :, 0, 0] - D[3][:, :, 0, 0])
            return D
        raise ValueError(op)


# ------------------------------------------------------------------ the loss
def samples(g, n=500, seed=17):
    """n seeded sample columns (i, j) and observations 1.85e-6 +- 1e-8."""
    rng = np.random.default_rng(seed)
    i = rng.integers(0, g.im, n)
    j = rng.integers(0, g.jm, n)
    obs = 1.85e-6 + 1e-8 * rng.standard_normal(n)
    return i, j, obs


def loss(rm, m, i, j, obs):
    """Sum over the samples of (column-average tracer-1 mixing ratio - obs)^2."""
    q = rm[i, j, :, 0].sum(axis=1) / m[i, j, :].sum(axis=1)
    return float(np.sum((q - obs) ** 2))


def loss_adjoint(rm, m, i, j, obs):
    """d loss / d rm (im, jm, lm, nt)."""
    M = m[i, j, :].sum(axis=1)
    q = rm[i, j, :, 0].sum(axis=1) / M
    lam = np.zeros_like(rm)
    np.add.at(lam[..., 0], (i, j), (2.0 * (q - obs) / M)[:, None] * np.ones(rm.shape[2])[None, :])
    return lam
