"""Synthetic meteorology for the TM5 transport step, regenerable bit-for-bit from a seed.

There is no meteo archive on the machines this runs on, so every input of a run is built here:
surface pressure, air mass from the hybrid a/b coefficients, a steady divergence-balanced
circulation (pu/pv) whose amplitude is modulated in time so that Check_CFL sometimes has to cut
the step, convective entrainment/detrainment, a Kzz profile, a CH4 emission field, CH4 loss
rates and the initial tracer.  Arrays are in the halo-free pydyn layout (C order, i first):

    m, ps-derived masses     (im, jm, lm)
    pu                       (im+1, jm, lm)   kg/s through x-faces, pu[0] == pu[im]
    pv                       (im, jm+1, lm)   kg/s through y-faces, zero at both poles
    entu detu entd detd      (im, jm, lm)     kg/s per cell, non-zero in the lowest CONV_LEVELS
    cloud_top cloud_lfs      (im, jm) int32   1-based levels, 0 = no cloud
    kvh                      (im, jm, lm-1)   m2/s at the interfaces l+1
    gph                      (im, jm, lm+1)   m, interface heights above the surface
    emis                     (im, jm)         kg CH4 per cell per month
    loss                     (im, jm, lm, 4)  1/s: OH, Cl, O(1D), soil
    rm                       (im, jm, lm, nt) kg tracer, slopes zero

The circulation is a stream function (rotational, no divergence on the discrete C-grid) plus a
divergent part whose layer weights sum to zero over the column, so the column mass is steady
and a 30-day run keeps its air mass (dynam0 re-derives the vertical flux from the horizontal
convergence).  Only IEEE +, -, *, / and the functions of `detmath` are used, never NumPy's
SIMD transcendentals or its pairwise sums, so the same seed gives the same bits on any x86 host.
"""

from dataclasses import dataclass, field

import numpy as np

from . import detmath

__all__ = ["GRAV", "AE", "XMCH4", "SEC_MONTH", "Grid", "Meteo", "grid", "CONFIGS"]

GRAV = 9.80665  # the model's g (balance.py's Binas value)
AE = 6.371e6  # earth radius, m
XMCH4 = 12.01115 + 1.0079 * 4.0  # g/mol, as the model's chem_param builds it
SEC_MONTH = 30.0 * 86400.0
CONV_LEVELS = 15  # convection is non-zero in the lowest 15 levels
NDYN_REF = 3600  # the step the flux amplitudes are scaled for
MOD_PERIOD = 4.0 * 86400.0  # period of the circulation's amplitude modulation, s
MOD_AMPLITUDE = 0.25


@dataclass
class Grid:
    """One global region: im x jm cells (dx, dy degrees), lm hybrid levels, convection over lmax."""

    name: str
    im: int
    jm: int
    lm: int
    lmax: int
    at: np.ndarray = field(repr=False, default=None)
    bt: np.ndarray = field(repr=False, default=None)
    dxyp: np.ndarray = field(repr=False, default=None)

    @property
    def dx(self):
        return 360.0 / self.im

    @property
    def dy(self):
        return 180.0 / self.jm

    def lat_faces(self):
        """Latitudes (radians) of the jm+1 y-faces, from the south pole."""
        rad = np.pi / 180.0
        return (-90.0 + self.dy * np.arange(self.jm + 1)) * rad

    def lat_centres(self):
        rad = np.pi / 180.0
        return (-90.0 + self.dy * (np.arange(self.jm) + 0.5)) * rad

    def lon_centres(self):
        rad = np.pi / 180.0
        return (-180.0 + self.dx * (np.arange(self.im) + 0.5)) * rad


def _hybrid(lm):
    """at (Pa), bt for the lm+1 interfaces, surface first: p = at + bt ps, 0 at the top."""
    eta = 1.0 - np.arange(lm + 1) / float(lm)
    bt = eta * eta
    at = 1.0e5 * (eta - bt)
    bt[0], at[0] = 1.0, 0.0
    bt[lm], at[lm] = 0.0, 0.0
    return at, bt


def grid(name):
    """glb6x4 (60 x 45 x 25, convection over 20 levels), glb3x2 (120 x 90 x 34, over 26), or the
    test grids tiny (12 x 8 x 6, over 5) and small (24 x 16 x 10, over 8)."""
    shapes = {
        "glb6x4": (60, 45, 25, 20),
        "glb3x2": (120, 90, 34, 26),
        "tiny": (12, 8, 6, 5),
        "small": (24, 16, 10, 8),
    }
    im, jm, lm, lmax = shapes[name]
    g = Grid(name, im, jm, lm, lmax)
    g.at, g.bt = _hybrid(lm)
    s = detmath.sin(g.lat_faces())
    g.dxyp = (g.dx * (np.pi / 180.0)) * (s[1:] - s[:-1]) * (AE * AE)
    return g


CONFIGS = {
    # name: (grid, tracers, simulated seconds)
    "c0": ("glb6x4", 1, 86400),
    "c1": ("glb3x2", 1, 30 * 86400),
}


def _seqsum(q, axis):
    q = np.moveaxis(q, axis, 0)
    acc = q[0].copy()
    for term in q[1:]:
        acc = acc + term
    return acc


def _normal(rng, shape):
    """An approximately normal deviate: the sum of 12 uniforms minus 6 (Irwin-Hall)."""
    acc = rng.uniform(0.0, 1.0, shape)
    for _ in range(11):
        acc = acc + rng.uniform(0.0, 1.0, shape)
    return acc - 6.0


def _smooth(rng, g, modes=3):
    """A smooth field on the cell centres (im, jm) scaled to max |value| = 1."""
    lon = g.lon_centres()[:, None]
    mu = detmath.sin(g.lat_centres())[None, :]
    acc = np.zeros((g.im, g.jm))
    for p in range(1, modes + 1):
        for q in range(0, 3):
            amp = rng.uniform(-1.0, 1.0)
            phase = rng.uniform(0.0, 2.0 * np.pi)
            leg = np.ones_like(mu)
            for _ in range(q):
                leg = leg * mu
            acc = acc + amp * detmath.cos(p * lon + phase) * leg
    return acc / np.max(np.abs(acc))


class Meteo:
    """Every input of a run on grid `g` from `seed` (see the module docstring)."""

    def __init__(self, g, seed=1, nt=1):
        self.g, self.seed, self.nt = g, seed, nt
        rng = np.random.default_rng(seed)
        im, jm, lm = g.im, g.jm, g.lm
        dt_ref = NDYN_REF / 2.0

        # surface pressure and air mass
        self.ps = 1.0e5 * (1.0 + 0.02 * _smooth(rng, g))
        phlb = g.at[None, None, :] + g.bt[None, None, :] * self.ps[..., None]
        self.phlb = phlb
        self.m = (phlb[..., :-1] - phlb[..., 1:]) * g.dxyp[None, :, None] / GRAV

        # rotational flow: stream function at the corners (i = 0..im periodic, j = 0..jm)
        mrow = _seqsum(self.m, 0) / float(im)  # (jm, lm)
        cx = 0.4 + 0.8 * rng.uniform(0.0, 1.0, lm)
        sign = np.where(rng.uniform(0.0, 1.0, lm) < 0.5, -1.0, 1.0)
        pu_zonal = (sign * cx)[None, :] * mrow / dt_ref  # (jm, lm)
        psi_z = np.zeros((jm + 1, lm))
        for j in range(jm):
            psi_z[j + 1] = psi_z[j] + pu_zonal[j]
        lon_f = (-180.0 + g.dx * np.arange(im + 1)) * (np.pi / 180.0)
        colat = detmath.sin(np.pi * np.arange(jm + 1) / float(jm))  # 0 at both poles
        pert = np.zeros((im + 1, jm + 1, lm))
        for k in (1, 2, 3):
            amp = rng.uniform(0.3, 1.0, lm)
            phase = rng.uniform(0.0, 2.0 * np.pi, lm)
            wave = detmath.cos(k * lon_f[:, None] + phase[None, :])  # (im+1, lm)
            envelope = colat * colat  # cos^2(lat): its meridional difference vanishes like cos(lat)
            for _ in range(k - 1):
                envelope = envelope * colat
            pert = pert + amp[None, None, :] * envelope[None, :, None] * wave[:, None, :]
        pert[im] = pert[0]
        # scale the perturbation so its meridional Courant number peaks at cy(l)
        pv_unit = pert[:-1] - pert[1:]  # (im, jm+1, lm)
        mlo = np.concatenate([self.m[:, :1], self.m], axis=1)
        mhi = np.concatenate([self.m, self.m[:, -1:]], axis=1)
        mmin = np.minimum(mlo, mhi)
        cy_unit = np.max(np.max(np.abs(pv_unit) * dt_ref / mmin, axis=0), axis=0)  # (lm,)
        cy = 0.4 + 0.75 * rng.uniform(0.0, 1.0, lm)
        bscale = cy / cy_unit
        psi = psi_z[None, :, :] + bscale[None, None, :] * pert  # (im+1, jm+1, lm)
        pu = psi[:, 1:, :] - psi[:, :-1, :]  # (im+1, jm, lm)
        pv = psi[:-1, :, :] - psi[1:, :, :]  # (im, jm+1, lm)
        pv[:, 0] = 0.0
        pv[:, jm] = 0.0

        # divergent part: layer weights summing to zero over the column
        chi = _smooth(rng, g, modes=2)
        w = np.zeros(lm)
        half = lm // 2
        w[:half] = 1.0
        w[half : 2 * half] = -1.0
        coslat_c = detmath.cos(g.lat_centres())
        coslat_f = detmath.cos(g.lat_faces())
        chi_w = np.concatenate([chi[-1:], chi], axis=0)  # periodic west neighbour
        dchi_x = chi_w[1:] - chi_w[:-1]  # (im, jm): chi(i) - chi(i-1)
        pud = np.zeros((im + 1, jm, lm))
        pud[:im] = (dchi_x * coslat_c[None, :])[..., None] * w[None, None, :]
        pud[im] = pud[0]
        pvd = np.zeros((im, jm + 1, lm))
        pvd[:, 1:jm] = ((chi[:, 1:] - chi[:, :-1]) * coslat_f[None, 1:jm])[..., None] * w[None, None, :]
        conv = (pud[:-1] - pud[1:]) + (pvd[:, :-1] - pvd[:, 1:])
        sd = np.zeros((im, jm, lm))
        acc = np.zeros((im, jm))
        for lev in range(lm - 1, 0, -1):
            acc = acc + conv[..., lev]
            sd[..., lev] = acc
        cz_unit = np.max(np.abs(sd) * dt_ref / self.m)
        dscale = 0.5 / cz_unit
        pu = pu + dscale * pud
        pv = pv + dscale * pvd
        # one factor on the whole flow so that the largest x Courant number is 1.2 at
        # amplitude 1 (1.5 at the modulation's peak): both parts keep their divergence
        mlo_x = np.concatenate([self.m[-1:], self.m], axis=0)
        mhi_x = np.concatenate([self.m, self.m[:1]], axis=0)
        cx_max = np.max(np.abs(pu) * dt_ref / np.minimum(mlo_x, mhi_x))
        total = 1.2 / cx_max
        self.pu0 = total * pu
        self.pv0 = total * pv
        self.pv0[:, 0] = 0.0
        self.pv0[:, jm] = 0.0
        self.pu0[im] = self.pu0[0]

        # convection: lognormal en/detrainment in the lowest CONV_LEVELS levels
        base = 0.02 * self.m / dt_ref
        shape = (im, jm, lm)
        low = (np.arange(lm) < CONV_LEVELS)[None, None, :]
        self.entu = np.where(low, base * detmath.exp(0.8 * _normal(rng, shape)), 0.0)
        self.detu = np.where(low, base * detmath.exp(0.8 * _normal(rng, shape)), 0.0)
        self.entd = np.where(low, 0.3 * base * detmath.exp(0.8 * _normal(rng, shape)), 0.0)
        self.detd = np.where(low, 0.3 * base * detmath.exp(0.8 * _normal(rng, shape)), 0.0)
        top = np.floor(rng.uniform(0.0, CONV_LEVELS + 1.0, (im, jm)))
        self.cloud_top = np.minimum(top, CONV_LEVELS).astype(np.int32)
        lfs = np.floor(rng.uniform(0.0, 1.0, (im, jm)) * (self.cloud_top + 1.0))
        self.cloud_lfs = np.minimum(lfs, self.cloud_top).astype(np.int32)

        # interface heights (isothermal hydrostatic column) and Kzz decaying with height
        scale_height = 287.0 * 260.0 / GRAV
        gph = np.zeros((im, jm, lm + 1))
        gph[..., 1:lm] = scale_height * detmath.log(self.ps[..., None] / phlb[..., 1:lm])
        gph[..., lm] = gph[..., lm - 1] + 2.0 * (gph[..., lm - 1] - gph[..., lm - 2])
        self.gph = gph
        decay = detmath.exp(-gph[..., 1:lm] / 1500.0)
        self.kvh = 1.0 + 99.0 * decay * rng.uniform(0.5, 1.0, (im, jm, lm - 1))

        # CH4 emission: lognormal flux with mean 1e-10 kg/m2/s, per cell per month
        flux = 1.0e-10 * detmath.exp(1.0 * _normal(rng, (im, jm)) - 0.5)
        self.emis = flux * g.dxyp[None, :] * SEC_MONTH

        # CH4 loss rates: OH 1e-6..1e-5 by latitude and level; Cl and O(1D) above level 25
        c2 = (coslat_c * coslat_c)[None, :, None]
        lev = np.arange(lm)[None, None, :].astype(np.float64)
        loss = np.zeros((im, jm, lm, 4))
        loss[..., 0] = 1.0e-6 + 9.0e-6 * c2 * (1.0 - lev / float(lm - 1)) * rng.uniform(0.9, 1.0, shape)
        strat = lev >= 25.0
        loss[..., 1] = np.where(strat, 2.0e-7 * (lev - 24.0), 0.0) * np.ones(shape)
        loss[..., 2] = np.where(strat, 5.0e-7 * (lev - 24.0), 0.0) * np.ones(shape)
        self.loss = loss

        # the tracer: 1.8e-6 mixing ratio with a 5 % Gaussian blob over North America
        rad = np.pi / 180.0
        lon = g.lon_centres()[:, None, None]
        lat = g.lat_centres()[None, :, None]
        rm = np.empty((im, jm, lm, nt))
        for t in range(nt):
            lon0, lat0 = (-100.0 + 40.0 * t) * rad, (40.0 - 10.0 * t) * rad
            dx, dy, dz = (lon - lon0) / (15.0 * rad), (lat - lat0) / (15.0 * rad), (lev - 3.0) / 5.0
            d2 = (dx * dx + dy * dy) + dz * dz
            q = 1.8e-6 * (1.0 + 0.05 * detmath.exp(-d2))
            rm[..., t] = q * self.m
        self.rm = rm

    # ------------------------------------------------------------------ time dependence
    def amplitude(self, t_sec):
        """The circulation's amplitude factor at time t_sec: a triangle wave between
        1 - MOD_AMPLITUDE and 1 + MOD_AMPLITUDE with period MOD_PERIOD, peaking at odd multiples
        of half the period.  Built from division, floor, abs, multiply and add so that a C or
        device copy (amplitude() in the native library) gives the same double."""
        x = float(t_sec) / MOD_PERIOD
        frac = x - float(np.floor(x))
        return 1.0 + MOD_AMPLITUDE * (1.0 - 4.0 * abs(frac - 0.5))

    def fluxes(self, t_sec):
        """(pu, pv) in kg/s at time t_sec: the steady circulation times amplitude(t_sec)."""
        a = self.amplitude(t_sec)
        return a * self.pu0, a * self.pv0

    def state(self):
        """The initial prognostic state: m and the four tracer arrays (slopes zero)."""
        z = np.zeros_like(self.rm)
        return {"m": self.m.copy(), "rm": self.rm.copy(), "rxm": z.copy(), "rym": z.copy(), "rzm": z.copy()}


# ------------------------------------------------------------------ nested (zoom) regions
@dataclass
class Child:
    """A nested region: im x jm cells over the parent cells ibeg..iend x jbeg..jend (1-based,
    inclusive), refined xref x yref in space and tref in time relative to the globe; the
    longitude/latitude bounds are the ones the region is declared with (the tree snaps them to
    the parent's cells)."""

    name: str
    im: int
    jm: int
    xbeg: float
    xend: float
    ybeg: float
    yend: float
    xref: int
    yref: int
    tref: int
    ibeg: int
    jbeg: int

    @property
    def iend(self):
        return self.ibeg + self.im // self.xref - 1

    @property
    def jend(self):
        return self.jbeg + self.jm // self.yref - 1


_TINY_CHILD = Child("tiny_child", 8, 8, -90.0, 30.0, -45.0, 45.0, 2, 2, 2, 4, 3)
_SMALL_CHILD = Child("small_child", 24, 18, -120.0, 60.0, -56.25, 45.0, 2, 2, 2, 5, 4)
# glb3x2 with a 1x1 degree child over 170W-50W, 10N-80N: parent cells 4..43 x 51..85
_NAM_CHILD = Child("nam1x1", 120, 70, -170.0, -50.0, 10.0, 80.0, 3, 2, 2, 4, 51)
# the reference's benchmarked tree's child (pydyn/gpu/bench.py:34-37): nam1x1 over 132W-60W,
# 20N-64N, 72 x 44, parent cells 17..40 x 56..77
_NAM72_CHILD = Child("nam1x1_72x44", 72, 44, -132.0, -60.0, 20.0, 64.0, 3, 2, 2, 17, 56)
# glb3x2's reduced grid as production runs it: rows (1-based) and their cluster sizes
GLB3X2_REDGRID = ((1, 2, 3, 4, 5, 6, 7, 84, 85, 86, 87, 88, 89, 90), (40, 8, 8, 4, 4, 4, 2, 2, 4, 4, 4, 8, 8, 40))

# parent grid name, seed, child, region 1's reduced grid (jred, clustsize) or None
ZOOMS = {
    "tiny2": ("tiny", 3, _TINY_CHILD, None),
    "small2": ("small", 5, _SMALL_CHILD, None),
    "c2": ("glb3x2", 1, _NAM_CHILD, None),
    "tiny2r": ("tiny", 3, _TINY_CHILD, ((1, 2, 8), (12, 4, 3))),
    "small2r": ("small", 5, _SMALL_CHILD, ((1, 2, 3, 14, 15, 16), (24, 8, 4, 4, 8, 24))),
    "c2r": ("glb3x2", 1, _NAM_CHILD, GLB3X2_REDGRID),
    # the tree the reference publishes its device timings on: glb3x2 on its reduced grid + nam1x1
    "pub": ("glb3x2", 1, _NAM72_CHILD, GLB3X2_REDGRID),
}


class NestedMeteo:
    """The inputs of a child region, refined from its parent's Meteo so that the two agree: the
    surface pressure, cloud levels, Kzz, heights and loss rates of the parent cell; extensive
    fields (en/detrainment, emission) split evenly over the cell's xref x yref children; air
    mass from the child's own cell areas; the tracer at the parent cell's mixing ratio.  The
    child's face fluxes interpolate the parent's linearly between the parent faces and share
    the cross-section (pu / yref, pv / xref), so every child cell's flux convergence is its
    parent cell's over xref yref and the vertical fluxes agree too."""

    def __init__(self, parent, child):
        p, c = parent, child
        g = p.g
        self.parent, self.c, self.nt = parent, child, parent.nt
        xr, yr = c.xref, c.yref
        im, jm, lm = c.im, c.jm, g.lm
        ip = (c.ibeg - 1) + np.arange(im) // xr  # parent cell of each child cell (0-based)
        jp = (c.jbeg - 1) + np.arange(jm) // yr
        cell = np.ix_(ip, jp)
        self.cell = cell
        rad = np.pi / 180.0
        lat_f = (-90.0 + g.dy * (c.jbeg - 1) + (g.dy / yr) * np.arange(jm + 1)) * rad
        s = detmath.sin(lat_f)
        self.dxyp = ((g.dx / xr) * rad) * (s[1:] - s[:-1]) * (AE * AE)
        self.ps = p.ps[cell]
        phlb = g.at[None, None, :] + g.bt[None, None, :] * self.ps[..., None]
        self.m = (phlb[..., :-1] - phlb[..., 1:]) * self.dxyp[None, :, None] / GRAV

        kx = np.arange(im + 1)
        f0, w = (c.ibeg - 1) + kx // xr, (kx % xr) / float(xr)
        pu = p.pu0[:, jp]  # (im_p+1, jm, lm)
        nxt = np.minimum(f0 + 1, g.im)
        self.pu0 = ((1.0 - w)[:, None, None] * pu[f0] + w[:, None, None] * pu[nxt]) / float(yr)
        ky = np.arange(jm + 1)
        g0, v = (c.jbeg - 1) + ky // yr, (ky % yr) / float(yr)
        pv = p.pv0[ip]  # (im, jm_p+1, lm)
        nxt = np.minimum(g0 + 1, g.jm)
        self.pv0 = ((1.0 - v)[None, :, None] * pv[:, g0] + v[None, :, None] * pv[:, nxt]) / float(xr)

        area = float(xr * yr)
        for k in ("entu", "detu", "entd", "detd"):
            setattr(self, k, getattr(p, k)[cell] / area)
        self.cloud_top = p.cloud_top[cell].copy()
        self.cloud_lfs = p.cloud_lfs[cell].copy()
        self.kvh, self.gph, self.loss = p.kvh[cell], p.gph[cell], p.loss[cell]
        self.emis = p.emis[cell] / area
        self.rm = (p.rm / p.m[..., None])[cell] * self.m[..., None]

    def amplitude(self, t_sec):
        return self.parent.amplitude(t_sec)

    def fluxes(self, t_sec):
        a = self.amplitude(t_sec)
        return a * self.pu0, a * self.pv0

    def state(self):
        z = np.zeros_like(self.rm)
        return {"m": self.m.copy(), "rm": self.rm.copy(), "rxm": z.copy(), "rym": z.copy(), "rzm": z.copy()}


def zoom(name, nt=1):
    """(parent Meteo, NestedMeteo) of a configuration in ZOOMS."""
    gname, seed, child, _ = ZOOMS[name]
    parent = Meteo(grid(gname), seed=seed, nt=nt)
    return parent, NestedMeteo(parent, child)


def redgrid(name):
    """Region 1's reduced grid (jred, clustsize) of a configuration in ZOOMS, or None."""
    return ZOOMS[name][3]


# ------------------------------------------------------------------ ensembles
class Ensemble:
    """`members` members of `tracers` tracers each (tracer t = member * tracers + k): member m's
    emission scaled by scale[m], drawn uniformly in [0.5, 1.5), into its first tracer only,
    and its initial tracers those of the meteo times (1 + 0.02 p_m), p_m a smooth field of
    max |p| = 1 (a child takes its parent cell's value)."""

    def __init__(self, parent, members, seed=11):
        rng = np.random.default_rng(seed)
        self.members, self.tracers = members, parent.nt
        self.scale = rng.uniform(0.5, 1.5, members)
        self.pert = [_smooth(rng, parent.g) for _ in range(members)]

    @property
    def nt(self):
        return self.members * self.tracers

    def subset(self, members):
        """The ensemble of the listed members only (their plain runs)."""
        out = object.__new__(Ensemble)
        out.members, out.tracers = len(members), self.tracers
        out.scale = self.scale[list(members)]
        out.pert = [self.pert[m] for m in members]
        return out

    def escale(self):
        """The emission scale of every tracer (zero: no emission)."""
        out = np.zeros(self.nt)
        out[:: self.tracers] = self.scale
        return out

    def state(self, met):
        """met.state() with the ensemble's tracers (im, jm, lm, members * tracers)."""
        st = met.state()
        rm = met.rm
        out = np.empty((*rm.shape[:3], self.nt))
        for m, p in enumerate(self.pert):
            field = p if not isinstance(met, NestedMeteo) else p[met.cell]
            factor = (1.0 + 0.02 * field)[:, :, None]
            for k in range(self.tracers):
                out[..., m * self.tracers + k] = rm[..., k] * factor
        st["rm"] = out
        for k in ("rxm", "rym", "rzm"):
            st[k] = np.zeros_like(out)
        return st
