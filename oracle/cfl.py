"""Check_CFL for one global region: the mass-only rehearsal that decides ndyn.

The rehearsal pushes the air mass alone through x, y, z, z, y, x (advectm_cfl.F90's split order
"xyzzyx" in blocks of three) with the sub-cycle caps (10, 1, 1); any operator over its cap fails
the interval, which is then retried with the next valid smaller ndyn (pydyn/cfl.py:98-103,
pydyn/zoom/check_cfl.py:79-290 with a tree of one region).
"""

import numpy as np

from .advection import closed_search, dynam0, periodic_search
from .lines import mass_after, nint

CAPS = (10, 1, 1)  # advectm_cfl.F90 cfl_max_nloop_{x,y,z}, pydyn/cfl.py:136


def new_valid_timestep(dtime, nread, cfl_outputstep, tref_max=1):
    """The largest dt < dtime dividing nread and cfl_outputstep that is a multiple of 2 tref_max
    (pydyn/cfl.py:208-213)."""
    for dt in range(dtime - 1, 0, -1):
        if nread % dt == 0 and cfl_outputstep % dt == 0 and dt % (2 * tref_max) == 0:
            return dt
    raise ValueError("new_valid_timestep: no valid time step")


def mass_x(m, am, max_nloop):
    """dynamum: every (j, l) row with its own nloop < max_nloop, am scaled and multiplied back
    by nloop (pydyn/cfl.py:140-167).  In place; False when a row exceeds the cap."""
    im, jm, lm = m.shape
    rows = jm * lm
    mm = m.reshape(im, rows).copy()
    a = am.reshape(im + 1, rows).copy()
    nloop = periodic_search(mm, a, max_nloop - 1)
    if nloop is None:
        return False
    for s in range(int(nloop.max())):
        sel = np.flatnonzero(nloop > s)
        mm[:, sel] = mass_after(mm[:, sel], a[:, sel])
    m[...] = mm.reshape(m.shape)
    am[...] = (a * nloop[None, :]).reshape(am.shape)
    return True


def mass_y(m, bm, max_nloop):
    """dynamvm: one nloop per layer, at most max_nloop (pydyn/cfl.py:170-183)."""
    for k in range(m.shape[2]):
        layer, b = np.swapaxes(m[:, :, k], 0, 1), np.swapaxes(bm[:, :, k], 0, 1)
        found = closed_search(layer, b, max_nloop, fresh=False)
        if found is None:
            return False
        for _ in range(found[0]):
            layer[...] = mass_after(layer, b)
        b[...] = b * found[0]
    return True


def mass_z(m, cm, max_nloop):
    """dynamwm: one n for the region, cm / n taken fresh, cm itself unchanged
    (pydyn/zoom/check_cfl.py:206-216 with pydyn/zoom/advect.py:415-441 on the whole region)."""
    mz, cz = np.moveaxis(m, 2, 0), np.moveaxis(cm, 2, 0)
    found = closed_search(mz, cz, max_nloop, fresh=True)
    if found is None:
        return False
    n, c = found
    c = cz / float(n)
    for _ in range(n):
        mz[...] = mass_after(mz, c)
    return True


def rehearse(st, caps=CAPS):
    """determine_cfl_iter twice: x, y, z then z, y, x on st's m, am, bm, cm, stopping at the
    first failure (pydyn/zoom/check_cfl.py:222-255).  Returns True when every operator passed."""
    ops = {
        "x": lambda: mass_x(st["m"], st["am"], caps[0]),
        "y": lambda: mass_y(st["m"], st["bm"], caps[1]),
        "z": lambda: mass_z(st["m"], st["cm"], caps[2]),
    }
    for op in "xyzzyx":
        if not ops[op]():
            return False
    return True


def check_cfl(st, fluxes_at, ndyn, n, nread, bt, tref_max=1, caps=CAPS, max_iteration=100):
    """Check_CFL with the CFL cache off (pydyn/zoom/check_cfl.py:258-290): for each of the n
    steps the fluxes of the step (fluxes_at(i, ndyn) -> pu, pv, i = 1..n), dynam0 and the
    rehearsal; m, am, bm, cm restored afterwards; on a failure ndyn goes to the next valid step,
    n = nint(n ndyn_old / ndyn) and am, bm, cm are scaled by ndyn / ndyn_old.
    Returns (ndyn, n, iterations)."""
    for iteration in range(1, max_iteration + 1):
        saved = {k: st[k].copy() for k in ("m", "am", "bm", "cm")}
        ok = True
        for i in range(1, n + 1):
            st["pu"], st["pv"] = fluxes_at(i, ndyn)
            st["am"], st["bm"], st["cm"] = dynam0(st["pu"], st["pv"], bt, ndyn / 2.0)
            ok = rehearse(st, caps)
            if not ok:
                break
        for k, v in saved.items():
            st[k] = v
        if ok:
            return ndyn, n, iteration
        ndyn_old = ndyn
        ndyn = new_valid_timestep(ndyn, nread, nread, tref_max)
        n = nint(float(n * ndyn_old) / float(ndyn))
        fraction = float(ndyn) / float(ndyn_old)
        for k in ("am", "bm", "cm"):
            st[k] = st[k] * fraction
    raise RuntimeError("check_cfl: exceeded maximum number of time steps")
