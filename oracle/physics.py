"""Kzz -> dkg, CH4 injection, CH4 decay and m2phlb1 for one global region (all cells owned)."""

import numpy as np

from .libm import exp as fdlibm_exp
from .lines import seqsum

GRAV = 9.80665  # pydyn/balance.py:40


def dkg(kvh, m, gph, lmax):
    """The exchange rate (im, jm, lmax) in kg/s across the interfaces l+1 for l < lmax-1, zero
    at lmax-1 (pydyn/diffusion.py:256-263)."""
    n = lmax - 1
    dz = gph[..., 2 : n + 2] - gph[..., :n]
    out = np.zeros((*m.shape[:2], lmax))
    out[..., :n] = np.maximum(0.0, kvh[..., :n]) * 2.0 * (m[..., 1 : n + 1] + m[..., :n]) / (dz * dz)
    return out


def month2dt(nsrce, sec_month, tref=1):
    """do_add_2d's month -> emission step factor (pydyn/emis.py:65-68)."""
    return (float(nsrce) / (2 * tref)) / sec_month


def inject(emis, factor, mol_mass, mol_mass_e, efrac=1.0):
    """The mass one do_add_2d call adds per cell (pydyn/emis.py:52-54)."""
    return ((emis * factor) * (mol_mass / mol_mass_e)) * efrac


def add_2d(rm, rzm, x, level, tracer):
    """rm += x, rzm -= x at the 1-based level and tracer, every cell (pydyn/emis.py:57-62,
    pydyn/zoom/physics.py:516-525)."""
    rm[:, :, level - 1, tracer - 1] = rm[:, :, level - 1, tracer - 1] + x
    rzm[:, :, level - 1, tracer - 1] = rzm[:, :, level - 1, tracer - 1] - x


def decay(loss, dt_sec):
    """exp(-(((L1 + L2) + L3) + L4) dt) per cell with the fdlibm exp (pydyn/chem.py:168-170)."""
    total = seqsum(np.moveaxis(loss, -1, 0))
    return fdlibm_exp(-(total * dt_sec))


def apply_decay(fields, factor, tracers):
    """rm, rxm, rym, rzm of the 1-based `tracers` times factor (pydyn/chem.py:173-178,
    pydyn/zoom/physics.py:528-539)."""
    for t in tracers:
        for q in fields:
            q[..., t - 1] = q[..., t - 1] * factor


def m2phlb1(m, dxyp, at, bt):
    """sp (im, jm) from the column mass and phlb (im, jm, lm+1) = at + bt sp
    (pydyn/zoom/massflow.py:282-290)."""
    p = seqsum(np.moveaxis(m, 2, 0)) * GRAV / dxyp[None, :]
    return p, at + bt * p[..., None]
