"""The zoom tree (TEST INFRASTRUCTURE): a global region with nested children, their two-way
exchange, the tree's Check_CFL and its driver -- a NumPy restatement of pydyn.zoom.

Arrays keep the Fortran bounds the model declares (pydyn/layout.py:26-48), since the exchange
is halo arithmetic: per region m (-1:im+2, -1:jm+2, 1:lm), rm..rzm (..., 1:nt), am bm cm
(0:im+1, 0:jm+1, 0:lm+1), pu (0:im, 1:jm, 1:lm), pv (1:im, 1:jm+1, 1:lm), sp (-1:im+2,
-1:jm+2, 1:1), phlb (-1:im+2, -1:jm+2, 1:lm+1), dxyp (1:jm).  `Arr` addresses them with
inclusive Fortran sections.  Region 1 runs the single-region operators of oracle.advection on
its interior views; a child runs the restated child bodies here.  Every function cites the
pydyn file:line it follows (pydyn/zoom/*.py).
"""

from dataclasses import dataclass

import numpy as np

from . import physics
from .advection import CFLError, advect_x, advect_y, advect_z, dynam0 as dynam0_fluxes
from .cfl import CAPS, mass_x, mass_y, new_valid_timestep
from .convection import convec
from .lines import CFL_ONE, face_fluxes, limit, mass_after, nint, pick_upwind, seqsum, shrink_progressive, slope_update

GRAV = physics.GRAV
SPLITORDER = "xyzvsccsvzyx"  # pydyn/zoom/tree.py:18
NOPS = len(SPLITORDER) // 2
CFL_SPLITORDER, CFL_NOPS = "xyzzyx", 3  # pydyn/zoom/check_cfl.py:40-41
FIELDS = ("m", "rm", "rxm", "rym", "rzm", "am", "bm", "cm", "pu", "pv", "sp", "phlb", "dxyp")


# ------------------------------------------------------------------ arrays with Fortran bounds
class Arr:
    """A NumPy array addressed by inclusive Fortran sections: a[lo:hi, j, ...] is the view
    (lo:hi, j, ...) of the declared bounds; axes left out are taken whole."""

    def __init__(self, data, lower):
        self.data, self.lower = data, tuple(lower)

    @classmethod
    def zeros(cls, *bounds):
        return cls(np.zeros(tuple(hi - lo + 1 for lo, hi in bounds)), [lo for lo, _ in bounds])

    def _ix(self, key):
        key = key if isinstance(key, tuple) else (key,)
        out = []
        for lo, k in zip(self.lower, key):
            if isinstance(k, slice):
                out.append(slice(None if k.start is None else k.start - lo, None if k.stop is None else k.stop - lo + 1))
            else:
                out.append(k - lo)
        return tuple(out)

    def __getitem__(self, key):
        return self.data[self._ix(key)]

    def __setitem__(self, key, value):
        self.data[self._ix(key)] = value

    def at(self, *index):
        """Elements at Fortran index arrays (broadcast)."""
        return self.data[tuple(np.asarray(ix) - lo for ix, lo in zip(index, self.lower))]

    def copy(self):
        return Arr(self.data.copy(), self.lower)


def declared(name, im, jm, lm, nt):
    """The bounds a region's array is declared with (pydyn/layout.py:26-36)."""
    b = {
        "m": ((-1, im + 2), (-1, jm + 2), (1, lm)),
        "am": ((0, im + 1), (0, jm + 1), (0, lm + 1)),
        "pu": ((0, im), (1, jm), (1, lm)),
        "pv": ((1, im), (1, jm + 1), (1, lm)),
        "sp": ((-1, im + 2), (-1, jm + 2), (1, 1)),
        "phlb": ((-1, im + 2), (-1, jm + 2), (1, lm + 1)),
        "dxyp": ((1, jm),),
    }
    if name in ("rm", "rxm", "rym", "rzm"):
        return (*b["m"], (1, nt))
    if name in ("bm", "cm"):
        return b["am"]
    return b[name]


# ------------------------------------------------------------------ the tree (initexit.F90)
@dataclass
class Region:
    """pydyn/zoom/tree.py:24-52: refinements are relative to the globe; parent 0 = the globe."""

    name: str
    im: int
    jm: int
    xbeg: float = -180.0
    xend: float = 180.0
    ybeg: float = -90.0
    yend: float = 90.0
    xref: int = 1
    yref: int = 1
    zref: int = 1
    tref: int = 1
    parent: int = 0
    xcyc: int = 0
    touch_sp: int = 0
    touch_np: int = 0
    redgrid: object = None  # RedGrid of region 1's x advection, or None


# ------------------------------------------------------------------ the reduced grid (redgrid.py)
class RedGrid:
    """Rows near the poles advected in x with clusters of `clustsize` cells merged into one
    (pydyn/redgrid.py:26-54): bands [(j 0-based, clustsize, imred)], imred per row."""

    def __init__(self, im, jm, jred, clustsize):
        self.im, self.jm = im, jm
        self.bands = []
        self.imred = np.full(jm, im, dtype=int)
        for j, c in zip(jred, clustsize):
            if im % c:
                raise ValueError(f"redgrid: im={im} not a multiple of clustsize {c}")
            self.bands.append((j - 1, c, im // c))
            self.imred[j - 1] = im // c

    @property
    def rows(self):
        return [b[0] for b in self.bands]


def _cluster_sums(row, n, c):
    """(n, ...) sums of consecutive clusters of c cells, accumulated in index order."""
    parts = row[: n * c].reshape(n, c, *row.shape[1:])
    acc = parts[:, 0].copy()
    for k in range(1, c):
        acc = acc + parts[:, k]
    return acc


def uni2red(grid, m, qs):
    """Clusters summed into cells 0..imred-1 of their row; the tracers beyond zeroed, m not
    (pydyn/redgrid.py:57-68)."""
    for j, c, n in grid.bands:
        m[:n, j] = _cluster_sums(m[:, j], n, c)
        for q in qs:
            q[:n, j] = _cluster_sums(q[:, j], n, c)
            q[n:, j] = 0.0


def uni2red_mass(grid, m):
    for j, c, n in grid.bands:
        m[:n, j] = _cluster_sums(m[:, j], n, c)


def uni2red_mf(grid, am):
    """am(i) = am(i clustsize) for i = 1..imred, am(0) = am(imred), the rest -1 (redgrid.py:71-77)."""
    for j, c, n in grid.bands:
        row = am[:, j]
        row[1 : n + 1] = row[c * np.arange(1, n + 1)]
        row[n + 1 :] = -1.0
        row[0] = row[n]


def red2uni(grid, m, qs, m_uni):
    """The clusters split again: tracers spread by m_uni / m of the cluster, m = m_uni on the
    reduced rows (redgrid.py:80-92)."""
    for j, c, n in grid.bands:
        scale = m_uni[: n * c, j].reshape(n, c, *m_uni.shape[2:]) / m[:n, j, None]
        for q in qs:
            q[: n * c, j] = (scale[..., None] * q[:n, j, None]).reshape(n * c, *q.shape[2:])
    for j, _, _ in grid.bands:
        m[:, j] = m_uni[:, j]


def red2uni_mass(grid, m, m_uni):
    for j, _, _ in grid.bands:
        m[:, j] = m_uni[:, j]


class Tree:
    """The region tree and what initexit derives from it (pydyn/zoom/tree.py:55-191): children,
    footprints ibeg..jend, interior scopes isr..jer, the zoomed/edge maps and each region's
    split order.  Regions count from 1; status[r] is the driver's position counter."""

    def __init__(self, regions, lm, at, bt, splitorder=SPLITORDER):
        self.reg = [None, *regions]
        self.n = len(regions)
        self.lm, self.at, self.bt = lm, np.asarray(at, dtype=float), np.asarray(bt, dtype=float)
        self.status = [0] * (self.n + 1)
        self._footprints()
        self.zoomed, self.edge = {}, {}
        for r in self.regions():
            self.zoomed[r], self.edge[r] = self._maps(r)
        self.order = self._orders(splitorder, NOPS)
        self._in_scope()

    def regions(self):
        return range(1, self.n + 1)

    def g(self, r):
        return self.reg[r] if r > 0 else Region("globe", self.reg[1].im, self.reg[1].jm)

    def ref(self, r, axis):
        """A refinement relative to the parent (tree.py:76-79)."""
        return getattr(self.g(r), axis + "ref") // getattr(self.g(self.g(r).parent), axis + "ref")

    def _footprints(self):  # tree.py:82-123
        n = self.n
        self.children = {r: [] for r in range(n + 1)}
        for r in range(2, n + 1):
            self.children[self.g(r).parent].append(r)
        g1 = self.g(1)
        dx = (g1.xend - g1.xbeg) / g1.im * g1.xref
        dy = (g1.yend - g1.ybeg) / g1.jm * g1.yref
        self.ibeg, self.iend, self.jbeg, self.jend, self.lbeg, self.lend = ({} for _ in range(6))
        for r in range(2, n + 1):
            g, p = self.g(r), self.g(self.g(r).parent)
            self.ibeg[r] = nint((g.xbeg - p.xbeg) / (dx / p.xref)) + 1
            self.jbeg[r] = nint((g.ybeg - p.ybeg) / (dy / p.yref)) + 1
            self.iend[r] = nint((g.xend - p.xbeg) / (dx / p.xref))
            self.jend[r] = nint((g.yend - p.ybeg) / (dy / p.yref))
            self.lbeg[r], self.lend[r] = 1, self.lm
            ncell = (self.iend[r] - self.ibeg[r] + 1) if self.iend[r] > self.ibeg[r] else (p.im - self.ibeg[r] + 1 + self.iend[r])
            if ncell * self.ref(r, "x") != g.im or (self.jend[r] - self.jbeg[r] + 1) * self.ref(r, "y") != g.jm:
                raise ValueError(f"region {r}: footprint and size disagree")
        self.isr, self.ier, self.jsr, self.jer = {1: 1}, {1: g1.im}, {1: 1}, {1: g1.jm}
        for r in range(2, n + 1):
            g, xr, yr = self.g(r), self.ref(r, "x"), self.ref(r, "y")
            self.isr[r], self.ier[r] = (1, g.im) if g.xcyc else (xr + 1, g.im - xr)
            self.jsr[r] = 1 if g.touch_sp == 1 else yr + 1
            self.jer[r] = g.jm if g.touch_np == 1 else g.jm - yr

    def _maps(self, r):  # blank_zoom_region, tree.py:126-159
        g = self.g(r)
        zoomed = np.full((g.im, g.jm), r, dtype=np.int32)
        edge = np.zeros((g.im, g.jm), dtype=np.int32)
        for c in self.children[r]:
            gc = self.g(c)
            ib, ie, jb, je = self.ibeg[c], self.iend[c], self.jbeg[c], self.jend[c]
            ii = np.arange(1, g.im + 1)[:, None]
            jj = np.arange(1, g.jm + 1)[None, :]
            inside_i = ((ii > ib) & (ii < ie)) | (gc.xcyc == 1)
            inside_j = ((jj > jb) | (gc.touch_sp == 1)) & ((jj < je) | (gc.touch_np == 1))
            zoomed[inside_i & inside_j] = c
            if gc.xcyc != 1:
                edge[ib - 1, jb - 1 : je] = 1
                edge[ie - 1, jb - 1 : je] = 1
            if gc.touch_sp != 1:
                edge[ib - 1 : ie, jb - 1] += 2
            if gc.touch_np != 1:
                edge[ib - 1 : ie, je - 1] += 2
        if g.parent != 0:
            xr, yr = self.ref(r, "x"), self.ref(r, "y")
            if g.xcyc != 1:
                edge[:xr, :] = -1
                edge[g.im - xr :, :] = -1
            if g.touch_sp != 1:
                edge[:, :yr] = -1
            if g.touch_np != 1:
                edge[:, g.jm - yr :] = -1
        return zoomed, edge

    def _orders(self, base, nops):  # define_splitorderzoom, tree.py:162-176
        order = {1: base}
        for r in range(2, self.n + 1):
            t = self.ref(r, "t")
            if t > 4 or t == 3:
                raise ValueError(f"region {r}: tref {t} not 1, 2 or 4")
            blocks = [order[self.g(r).parent][k : k + nops] for k in range(0, len(order[self.g(r).parent]), nops)]
            order[r] = "".join(b if i % 2 == 0 else b[::-1] for b in blocks for i in range(t))
        return order

    def _in_scope(self):  # tree.py:178-191
        g1 = self.g(1)
        if not (g1.xcyc == 1 and g1.touch_sp == 1 and g1.touch_np == 1):
            raise NotImplementedError("region 1 must be global")
        for r in range(2, self.n + 1):
            g, p = self.g(r), self.g(self.g(r).parent)
            if g.xcyc or g.touch_sp or g.touch_np:
                raise NotImplementedError("cyclic or pole-touching children")
            if not (1 < self.ibeg[r] < self.iend[r] < p.im and 1 < self.jbeg[r] and self.jend[r] < p.jm):
                raise NotImplementedError("a child must lie strictly inside its parent")


def new_state(tree, nt):
    """{region: {name: Arr}} of zeros with the declared bounds (pydyn/zoom/state.py:30-37)."""
    st = {}
    for r in tree.regions():
        g = tree.g(r)
        st[r] = {k: Arr.zeros(*declared(k, g.im, g.jm, tree.lm, nt)) for k in FIELDS}
    return st


def ntracers(st):
    return st[1]["rm"].data.shape[3]


# ------------------------------------------------------------------ exchange (zoom/exchange.py)
def put_xedges(tree, st, r, tracers=True):
    """The parent's cells next to and on each x-wall of every child into the child's halo and
    interface cells, scaled to the child's cell size (exchange.py:17-60)."""
    f = st[r]
    for c in tree.children[r]:
        gc = tree.g(c)
        xr, yr, zr = tree.ref(c, "x"), tree.ref(c, "y"), tree.ref(c, "z")
        if gc.im // gc.xref == tree.g(1).im:
            continue
        fc, jmc, lm = st[c], gc.jm, tree.lm
        yz, xyz = 1.0 / (yr * zr), 1.0 / (xr * yr * zr)
        jp = (tree.jbeg[c] + np.arange(jmc) // yr)[:, None]
        lp = (tree.lbeg[c] + np.arange(lm) // zr)[None, :]
        ib, ie = tree.ibeg[c], tree.iend[c]
        names = [("m", None, xyz)]
        if tracers:
            names += [("rm", True, xyz), ("rxm", True, yz), ("rym", True, yz), ("rzm", True, yz)]
        for name, per_tracer, inner in names:
            src, dst = f[name], fc[name]
            for t in range(1, ntracers(st) + 1) if per_tracer else [None]:
                ext = () if t is None else (t,)
                west, west_in = src.at(ib - 1, jp, lp, *ext) * yz, src.at(ib, jp, lp, *ext) * inner
                east_in, east = src.at(ie, jp, lp, *ext) * inner, src.at(ie + 1, jp, lp, *ext) * yz
                dst[(0, slice(1, jmc), slice(1, lm), *ext)] = west
                dst[(gc.im + 1, slice(1, jmc), slice(1, lm), *ext)] = east
                for i in range(1, xr + 1):
                    dst[(i, slice(1, jmc), slice(1, lm), *ext)] = west_in
                    dst[(gc.im + 1 - i, slice(1, jmc), slice(1, lm), *ext)] = east_in


def put_yedges(tree, st, r, tracers=True):
    """The same across the y-walls (exchange.py:63-105)."""
    f, imp = st[r], tree.g(r).im
    for c in tree.children[r]:
        gc = tree.g(c)
        xr, yr, zr = tree.ref(c, "x"), tree.ref(c, "y"), tree.ref(c, "z")
        fc, imc, lm = st[c], gc.im, tree.lm
        xz, xyz = 1.0 / (xr * zr), 1.0 / (xr * yr * zr)
        ip = (np.mod(tree.ibeg[c] - 1 + np.arange(imc) // xr, imp) + 1)[:, None]
        lp = (tree.lbeg[c] + np.arange(lm) // zr)[None, :]
        jb, je = tree.jbeg[c], tree.jend[c]
        names = [("m", None, xyz)]
        if tracers:
            names += [("rm", True, xyz), ("rxm", True, xz), ("rym", True, xz), ("rzm", True, xz)]
        for name, per_tracer, inner in names:
            src, dst = f[name], fc[name]
            for t in range(1, ntracers(st) + 1) if per_tracer else [None]:
                ext = () if t is None else (t,)
                south, south_in = src.at(ip, jb - 1, lp, *ext) * xz, src.at(ip, jb, lp, *ext) * inner
                north_in, north = src.at(ip, je, lp, *ext) * inner, src.at(ip, je + 1, lp, *ext) * xz
                dst[(slice(1, imc), 0, slice(1, lm), *ext)] = south
                dst[(slice(1, imc), gc.jm + 1, slice(1, lm), *ext)] = north
                for j in range(1, yr + 1):
                    dst[(slice(1, imc), j, slice(1, lm), *ext)] = south_in
                    dst[(slice(1, imc), gc.jm + 1 - j, slice(1, lm), *ext)] = north_in


def _xedges_apply(tree, r):
    xr, g = tree.ref(r, "x"), tree.g(r)
    return not (xr == 1 or g.im // g.xref == tree.g(1).im)


def compress_xedges(tree, st, r, js, je, ls, le, tracers=True):
    """The interface cells 1..xref and im-xref+1..im summed into the last / first of them, the
    halo moved next to them (exchange.py:109-129)."""
    if not _xedges_apply(tree, r):
        return
    f, xr, im = st[r], tree.ref(r, "x"), tree.g(r).im
    rows = (slice(js, je), slice(ls, le))
    names = ("m", "rm") if tracers else ("m",)
    for name in names:
        q = f[name]
        q[(xr, *rows)] = seqsum(q[(slice(1, xr), *rows)])
        q[(im - xr + 1, *rows)] = seqsum(q[(slice(im - xr + 1, im), *rows)])
    for name in ("m", "rm", "rxm", "rym", "rzm") if tracers else ("m",):
        q = f[name]
        q[(xr - 1, *rows)] = q[(0, *rows)]
        q[(im - xr + 2, *rows)] = q[(im + 1, *rows)]


def uncompress_xedges(tree, st, r, js, je, ls, le, tracers=True):
    """The edge cell's value shared evenly over its interface cells again (exchange.py:132-144)."""
    if not _xedges_apply(tree, r):
        return
    f, xr, im = st[r], tree.ref(r, "x"), tree.g(r).im
    rows = (slice(js, je), slice(ls, le))
    for lo, hi, e in ((1, xr, xr), (im - xr + 1, im, im - xr + 1)):
        for name in ("m", "rm", "rxm", "rym", "rzm") if tracers else ("m",):
            q = f[name]
            share = q[(e, *rows)] / xr
            q[(slice(lo, hi), *rows)] = share[None]


def compress_yedges(tree, st, r, is_, ie, ls, le, tracers=True):
    """compress in y (exchange.py:147-165)."""
    yr, jm = tree.ref(r, "y"), tree.g(r).jm
    if yr == 1 or r == 1:
        return
    f = st[r]
    cols, lev = slice(is_, ie), slice(ls, le)
    for name in ("m", "rm") if tracers else ("m",):
        q = f[name]
        q[cols, yr, lev] = seqsum(np.moveaxis(q[cols, slice(1, yr), lev], 1, 0))
        q[cols, jm - yr + 1, lev] = seqsum(np.moveaxis(q[cols, slice(jm - yr + 1, jm), lev], 1, 0))
    for name in ("m", "rm", "rxm", "rym", "rzm") if tracers else ("m",):
        q = f[name]
        q[cols, yr - 1, lev] = q[cols, 0, lev]
        q[cols, jm - yr + 2, lev] = q[cols, jm + 1, lev]


def uncompress_yedges(tree, st, r, is_, ie, ls, le, tracers=True):
    """uncompress in y (exchange.py:168-177)."""
    yr, jm = tree.ref(r, "y"), tree.g(r).jm
    if yr == 1 or r == 1:
        return
    f = st[r]
    cols, lev = slice(is_, ie), slice(ls, le)
    for lo, hi, e in ((1, yr, yr), (jm - yr + 1, jm, jm - yr + 1)):
        for name in ("m", "rm", "rxm", "rym", "rzm") if tracers else ("m",):
            q = f[name]
            share = q[cols, e, lev] / yr
            q[cols, slice(lo, hi), lev] = share[:, None]


def mix_edges(tree, st, r, tracers=True):
    """Every parent-cell block of the child's ring (edge == -1) replaced by its mean, j-outer /
    i-inner accumulation (exchange.py:181-231)."""
    g = tree.g(r)
    xr, yr, zr = tree.ref(r, "x"), tree.ref(r, "y"), tree.ref(r, "z")
    if zr != 1:
        raise ValueError("mix_edges: zooming in z direction not supported!")
    xedges = _xedges_apply(tree, r)
    if not xedges and r == 1:
        return
    ysp, ynp = g.touch_sp != 1, g.touch_np != 1
    im, jm = g.im, g.jm
    imp, jmp, xyr = im // xr, jm // yr, xr * yr
    ring = tree.edge[r] == -1

    def means(q):
        part = np.where(ring, q / xyr, 0.0).reshape(imp, xr, jmp, yr)
        acc = np.zeros((imp, jmp))
        for b in range(yr):
            for a in range(xr):
                acc = acc + part[:, a, :, b]
        return acc

    def spread(q, blk):
        ip, jp = np.arange(im) // xr, np.arange(jm) // yr
        if xedges:
            q[:xr, :] = blk[0, jp][None, :]
            q[im - xr :, :] = blk[imp - 1, jp][None, :]
        if ysp:
            q[:, :yr] = blk[ip, 0][:, None]
        if ynp:
            q[:, jm - yr :] = blk[ip, jmp - 1][:, None]

    f = st[r]
    for lev in range(1, tree.lm + 1):
        q = f["m"][1:im, 1:jm, lev]
        spread(q, means(q))
    if tracers:
        for t in range(1, ntracers(st) + 1):
            for lev in range(1, tree.lm + 1):
                qs = [f[k][1:im, 1:jm, lev, t] for k in ("rm", "rxm", "rym", "rzm")]
                blocks = [means(q) for q in qs]
                for q, b in zip(qs, blocks):
                    spread(q, b)


def _block_sums_ij(q, imp, xr, jmp, yr):
    """(im, jm, ...) -> (imp, jmp, ...) accumulated i-outer / j-inner (exchange.py:244-249)."""
    part = q.reshape(imp, xr, jmp, yr, *q.shape[2:])
    acc = np.zeros((imp, jmp, *q.shape[2:]))
    for a in range(xr):
        for b in range(yr):
            acc = acc + part[:, a, :, b]
    return acc


def update_parent(tree, st, r, tracers=True):
    """The parent's m and rm over the footprint from the child's block sums; its slopes as they
    were (exchange.py:235-258)."""
    if r == 1:
        return
    g, p = tree.g(r), tree.g(r).parent
    xr, yr, zr = tree.ref(r, "x"), tree.ref(r, "y"), tree.ref(r, "z")
    imp, jmp, lm = g.im // xr, g.jm // yr, tree.lm
    ib, ie, jb, je = tree.ibeg[r], tree.iend[r], tree.jbeg[r], tree.jend[r]
    f, fp = st[r], st[p]
    fp["m"][ib:ie, jb:je, 1:lm] = _block_sums_ij(f["m"][1 : g.im, 1 : g.jm, 1:lm], imp, xr, jmp, yr)
    if zr != 1:
        raise ValueError("update_parent: zooming in z direction not supported")
    if tracers:
        for t in range(1, ntracers(st) + 1):
            fp["rm"][ib:ie, jb:je, 1:lm, t] = _block_sums_ij(f["rm"][1 : g.im, 1 : g.jm, 1:lm, t], imp, xr, jmp, yr)


# ------------------------------------------------------------------ mass flow (zoom/massflow.py)
def _column_pressure(tree, st, r):
    g, f = tree.g(r), st[r]
    return seqsum(np.moveaxis(f["m"][1 : g.im, 1 : g.jm, 1 : tree.lm], 2, 0)) * GRAV / f["dxyp"][1 : g.jm][None, :]


def _store_pressure(tree, st, r, p):
    g, f = tree.g(r), st[r]
    f["sp"].data[...] = 0.0
    f["sp"][1 : g.im, 1 : g.jm, 1] = p
    f["phlb"][1 : g.im, 1 : g.jm, 1 : tree.lm + 1] = tree.at + tree.bt * p[..., None]


def m2phlb1(tree, st, r=1):
    """sp from the column mass, phlb = at + bt sp, sp's halo zeroed (massflow.py:24-32)."""
    _store_pressure(tree, st, r, _column_pressure(tree, st, r))


def m2phlb(tree, st, r):
    """m2phlb1 for a child, the ring cells' pressure rescaled to the parent cell's area
    (massflow.py:35-58)."""
    g, f = tree.g(r), st[r]
    xr, yr = tree.ref(r, "x"), tree.ref(r, "y")
    dxyp, dxyp_p = f["dxyp"], st[g.parent]["dxyp"]
    p = _column_pressure(tree, st, r)
    for j in range(1, g.jm + 1):
        for i in range(1, g.im + 1):
            inner_i, inner_j = xr < i < g.im - xr + 1, yr < j < g.jm - yr + 1
            if (inner_i and inner_j) or (g.xcyc == 1 and inner_j):
                continue
            if (g.touch_np == 1 and j > g.jm - yr) or (g.touch_sp == 1 and j < yr + 1):
                continue
            jpar = tree.jbeg[r] + (j - 1) // yr
            p[i - 1, j - 1] = p[i - 1, j - 1] * dxyp[j] * xr * yr / dxyp_p[jpar]
    _store_pressure(tree, st, r, p)


def coarsen_block(tree, st, r):
    """The parent's am, bm, cm and m over the footprint rebuilt from the child's, fluxes times
    tref (massflow.py:79-147)."""
    p = tree.g(r).parent
    g, f, fp = tree.g(r), st[r], st[p]
    xr, yr, zr, tr = (tree.ref(r, a) for a in "xyzt")
    imp, jmp, lm = g.im // xr, g.jm // yr, tree.lm
    ib, jb, lb = tree.ibeg[r], tree.jbeg[r], tree.lbeg[r]
    if zr != 1:
        raise ValueError("coarsen_region: zooming in z direction not supported")
    ipg = np.mod(ib - 2 + np.arange(imp + 1), tree.g(p).im) + 1  # parent face of child face block 0..imp
    jpg = jb + np.arange(jmp)
    lpg = lb + np.arange(lm)
    # am: child faces xr*ip summed over each block's rows (j ascending)
    faces = f["am"][0 : g.im, 1 : g.jm, 1:lm][xr * np.arange(imp + 1)].reshape(imp + 1, jmp, yr, lm)
    acc = np.zeros((imp + 1, jmp, lm))
    for b in range(yr):
        acc = acc + faces[:, :, b] * tr
    fp["am"].data[np.ix_(ipg - fp["am"].lower[0], jpg - fp["am"].lower[1], lpg - fp["am"].lower[2])] = acc
    # bm: child faces 1 + (jp-1) yr summed over each block's columns (i ascending)
    faces = f["bm"][1 : g.im, 1 : g.jm + 1, 1:lm][:, yr * np.arange(jmp + 1)].reshape(imp, xr, jmp + 1, lm)
    acc = np.zeros((imp, jmp + 1, lm))
    for a in range(xr):
        acc = acc + faces[:, a] * tr
    bm = fp["bm"]
    bm.data[np.ix_(ipg[1:] - bm.lower[0], jb - 1 + np.arange(1, jmp + 2) - bm.lower[1], lpg - bm.lower[2])] = acc
    # cm: summed over the block, i outer / j inner
    cells = f["cm"][1 : g.im, 1 : g.jm, 1:lm].reshape(imp, xr, jmp, yr, lm)
    acc = np.zeros((imp, jmp, lm))
    for a in range(xr):
        for b in range(yr):
            acc = acc + cells[:, a, :, b] * tr
    cm = fp["cm"]
    cm.data[np.ix_(ipg[1:] - cm.lower[0], jpg - cm.lower[1], lpg - cm.lower[2])] = acc
    # m: summed over the block, j outer / i inner
    cells = f["m"][1 : g.im, 1 : g.jm, 1:lm].reshape(imp, xr, jmp, yr, lm)
    acc = np.zeros((imp, jmp, lm))
    for b in range(yr):
        for a in range(xr):
            acc = acc + cells[:, a, :, b]
    m = fp["m"]
    m.data[np.ix_(ipg[1:] - m.lower[0], jpg - m.lower[1], lpg - m.lower[2])] = acc


def coarsen_region(tree, st, r):
    """coarsen_block into the parent, the parent's m2phlb, recursively up (massflow.py:62-76)."""
    p = tree.g(r).parent
    if p <= 0:
        return
    coarsen_block(tree, st, r)
    if p != 1:
        m2phlb(tree, st, p)
    else:
        m2phlb1(tree, st, p)
    coarsen_region(tree, st, p)


def coarsen_ambm(tree, st, r):
    """Each child's outer-interface fluxes from the parent's face flux divided over the child
    faces it covers (massflow.py:150-177)."""
    f = st[r]
    for c in tree.children[r]:
        gc, fc = tree.g(c), st[c]
        xr, yr, zr = tree.ref(c, "x"), tree.ref(c, "y"), tree.ref(c, "z")
        ib, ie, jb, je, lb, le = tree.ibeg[c], tree.iend[c], tree.jbeg[c], tree.jend[c], tree.lbeg[c], tree.lend[c]
        if gc.im // gc.xref != tree.g(1).im:
            for j in range(jb, je + 1):
                for lev in range(lb, le + 1):
                    jc, lc = (j - jb) * yr + 1, (lev - lb) * zr + 1
                    fc["am"][xr - 1, jc : jc + yr - 1, lc : lc + zr - 1] = f["am"][ib - 1, j, lev] / (yr * zr)
                    fc["am"][gc.im - xr + 1, jc : jc + yr - 1, lc : lc + zr - 1] = f["am"][ie, j, lev] / (yr * zr)
        for i in range(ib, ie + 1):
            for lev in range(lb, le + 1):
                ic, lc = (i - ib) * xr + 1, (lev - lb) * zr + 1
                if gc.touch_sp != 1:
                    fc["bm"][ic : ic + xr - 1, yr, lc : lc + zr - 1] = f["bm"][i, jb, lev] / (xr * zr)
                if gc.touch_np != 1:
                    fc["bm"][ic : ic + xr - 1, gc.jm - yr + 2, lc : lc + zr - 1] = f["bm"][i, je + 1, lev] / (xr * zr)


def dynam0(tree, st, r, ndyn):
    """dynam0 of a region: dt = ndyn / (2 tref), am bm cm zeroed whole first, a cyclic region's
    periodic halo faces (massflow.py:185-203).  The fluxes are kernels.py:102-125's, which keep
    pu's face 0 (the single-region dynam0 of oracle.advection ties it to face im)."""
    g, f, lm = tree.g(r), st[r], tree.lm
    im, jm = g.im, g.jm
    dt = ndyn / (2.0 * g.tref)
    pu = f["pu"][0:im, 1:jm, 1:lm].copy()
    am, bm, cm = dynam0_fluxes(pu, f["pv"][1:im, 1 : jm + 1, 1:lm].copy(), tree.bt, dt)
    am[0] = dt * pu[0]
    for k in ("am", "bm", "cm"):
        f[k].data[...] = 0.0
    f["am"][0:im, 1:jm, 1:lm] = am
    f["bm"][1:im, 1 : jm + 1, 1:lm] = bm
    f["cm"][1:im, 1:jm, 1 : lm - 1] = cm[..., 1:lm]
    if g.xcyc == 1:
        a = f["am"]
        a[0, :, :] = a[im, :, :].copy()
        a[im + 1, :, :] = a[1, :, :].copy()


def setup_massflow(tree, st, ndyn):
    """dynam0 and coarsen_region region by region, then coarsen_ambm (massflow.py:181-193)."""
    for r in tree.regions():
        dynam0(tree, st, r, ndyn)
        coarsen_region(tree, st, r)
    for r in tree.regions():
        coarsen_ambm(tree, st, r)


# ------------------------------------------------------------------ advection (zoom/advect.py)
def block(tree, r, status=None, nops=NOPS, order=None):
    """The operators of the current block and whether this is the second half of the parent
    step (advect.py:47-59)."""
    t = tree.ref(r, "t")
    s = tree.status[r] if status is None else status
    order = tree.order[r] if order is None else order
    start = (s // (nops * t)) * nops * t
    return order[start : start + nops], s % (nops * t) >= nops


def x_scope(tree, r, blk):  # advect.py:62-78
    g = tree.g(r)
    yr = 0 if r == 1 else tree.ref(r, "y")
    x_seen, js, je = False, None, None
    for d in blk:
        if d == "x":
            x_seen = True
        elif d == "y":
            if not x_seen:
                js, je = 1, g.jm
            else:
                js = 1 if g.touch_sp == 1 else yr + 1
                je = g.jm if g.touch_np == 1 else g.jm - yr
    return js, je


def y_scope(tree, r, blk):  # advect.py:81-96
    g = tree.g(r)
    xr = 0 if r == 1 else tree.ref(r, "x")
    y_seen, is_, ie = False, None, None
    for d in blk:
        if d == "x":
            is_, ie = (1, g.im) if (not y_seen or g.xcyc == 1) else (xr + 1, g.im - xr)
        elif d == "y":
            y_seen = True
    return is_, ie


def z_scope(tree, r, blk):  # advect.py:99-118
    g = tree.g(r)
    xr, yr = (0, 0) if r == 1 else (tree.ref(r, "x"), tree.ref(r, "y"))
    z_seen = False
    is_ = ie = js = je = None
    for d in blk:
        if d == "x":
            is_, ie = (1, g.im) if (not z_seen or g.xcyc == 1) else (xr + 1, g.im - xr)
        elif d == "y":
            if not z_seen:
                js, je = 1, g.jm
            else:
                js = 1 if g.touch_sp == 1 else yr + 1
                je = g.jm if g.touch_np == 1 else g.jm - yr
        elif d == "z":
            z_seen = True
    return is_, ie, js, je


def _outer_x(tree, st, r, second):
    """The child's outer x interface faces to zero in the second half (advect.py:121-136):
    the saved slabs, or None."""
    g, xr = tree.g(r), tree.ref(r, "x")
    if not (second and g.im // g.xref < tree.g(1).im):
        return None
    am = st[r]["am"]
    saved = (am[xr - 1, :, :].copy(), am[g.im - xr + 1, :, :].copy())
    am[xr - 1, :, :] = 0.0
    am[g.im - xr + 1, :, :] = 0.0
    return saved


def _restore_outer_x(tree, st, r, saved):
    if saved is not None:
        g, xr, am = tree.g(r), tree.ref(r, "x"), st[r]["am"]
        am[xr - 1, :, :], am[g.im - xr + 1, :, :] = saved


def _outer_y(tree, st, r, second):  # advect.py:139-157
    g, yr, bm = tree.g(r), tree.ref(r, "y"), st[r]["bm"]
    saved = [None, None]
    if second:
        if g.touch_sp != 1:
            saved[0] = bm[:, yr, :].copy()
            bm[:, yr, :] = 0.0
        if g.touch_np != 1:
            saved[1] = bm[:, g.jm - yr + 2, :].copy()
            bm[:, g.jm - yr + 2, :] = 0.0
    return saved


def _restore_outer_y(tree, st, r, saved):
    g, yr, bm = tree.g(r), tree.ref(r, "y"), st[r]["bm"]
    if saved[0] is not None:
        bm[:, yr, :] = saved[0]
    if saved[1] is not None:
        bm[:, g.jm - yr + 2, :] = saved[1]


def interior(f, im, jm, lm):
    """Region-1 views of m, rm..rzm in the halo-free layout (state.py:47-53)."""
    return (f["m"][1:im, 1:jm, 1:lm], *(f[k][1:im, 1:jm, 1:lm] for k in ("rm", "rxm", "rym", "rzm")))


def advect_x_zoom(tree, st, r, limits=True):
    """put_xedges, then dynamu on the block's scope, the outer faces zeroed in the second half
    (advect.py:160-176)."""
    put_xedges(tree, st, r)
    blk, second = block(tree, r)
    js, je = x_scope(tree, r, blk)
    saved = _outer_x(tree, st, r, second)
    if r == 1:
        x_region1(tree, st, js, je, limits)
    else:
        x_child(tree, st, r, js, je, 1, tree.lm, limits)
        uncompress_xedges(tree, st, r, js, je, 1, tree.lm)
        mix_edges(tree, st, r)
    _restore_outer_x(tree, st, r, saved)


def advect_y_zoom(tree, st, r, limits=True):
    """put_yedges, then dynamv on its scope (advect.py:179-194)."""
    put_yedges(tree, st, r)
    blk, second = block(tree, r)
    is_, ie = y_scope(tree, r, blk)
    saved = _outer_y(tree, st, r, second)
    if r == 1:
        y_region1(tree, st, limits)
    else:
        y_child(tree, st, r, is_, ie, 1, tree.lm, limits)
        uncompress_yedges(tree, st, r, is_, ie, 1, tree.lm)
        mix_edges(tree, st, r)
    _restore_outer_y(tree, st, r, saved)


def advect_z_zoom(tree, st, r, limits=True):
    """dynamw on its scope (advect.py:197-206)."""
    blk, _ = block(tree, r)
    is_, ie, js, je = z_scope(tree, r, blk)
    if r == 1:
        g, f, lm = tree.g(1), st[1], tree.lm
        advect_z(*interior(f, g.im, g.jm, lm), f["cm"][1 : g.im, 1 : g.jm, 0:lm], limits)
    else:
        z_child(tree, st, r, is_, ie, js, je, limits)
        mix_edges(tree, st, r)


def x_region1(tree, st, js, je, limits):
    """dynamu(1) on the interior rows (on the reduced grid when region 1 has one), then the
    periodic x halos its closing loop writes for rows js..je -- in the reduced representation,
    before the clusters are split -- and, for every reduced row, q(0) = q(im) after the split
    (pydyn/zoom/advect.py:253-297)."""
    g, f, lm = tree.g(1), st[1], tree.lm
    im, jm = g.im, g.jm
    fields = interior(f, im, jm, lm)
    am = f["am"][0:im, 1:jm, 1:lm]
    grid = g.redgrid
    if grid is None:
        advect_x(*fields, am, limits)
    else:
        m_uni, am_uni = fields[0].copy(), am.copy()
        uni2red(grid, fields[0], fields[1:])
        uni2red_mf(grid, am)
        for j in range(jm):
            n = grid.imred[j]
            advect_x(*(q[:n, j : j + 1] for q in fields), am[: n + 1, j : j + 1], limits)
    for j in range(js, je + 1):
        n = grid.imred[j - 1] if grid is not None else im
        for name in ("rm", "rxm", "rym", "rzm", "m"):
            q = f[name]
            q[0, j] = q[n, j].copy()
            q[n + 1, j] = q[1, j].copy()
    if grid is not None:
        red2uni(grid, fields[0], fields[1:], (m_uni + am_uni[:-1]) - am_uni[1:])
        am[...] = am_uni
        for j0 in grid.rows:
            for name in ("rm", "rxm", "rym", "rzm", "m"):
                q = f[name]
                q[0, j0 + 1] = q[im, j0 + 1].copy()


def y_region1(tree, st, limits):
    """dynamv(1) on the interior, and the layer's faces outside the interior scaled as dynamv
    scales the whole layer (advect.py:209-239)."""
    g, f, lm = tree.g(1), st[1], tree.lm
    im, jm = g.im, g.jm
    layers = advect_y(*interior(f, im, jm, lm), f["bm"][1:im, 1 : jm + 1, 1:lm], limits)
    layer_halo(f["bm"].data, im, jm, layers)


def layer_halo(bm, im, jm, layers):
    """Region 1's bm faces outside bm(1:im, 1:jm+1, l) (data index = Fortran index, lower bound
    0) through dynamv's whole-layer scaling: per layer l+1 that sub-cycled, (n-1)/n for n = 2..
    nloop, then times nloop (advect.py:212-226)."""
    for lev, nl in enumerate(layers, start=1):
        if nl == 1:
            continue
        layer = bm[:, :, lev]
        outside = np.ones(layer.shape, dtype=bool)
        outside[1 : im + 1, 1 : jm + 2] = False
        halo = layer[outside]
        for n in range(2, int(nl) + 1):
            halo = shrink_progressive(halo, n)
        layer[outside] = halo * int(nl)


def _row_search(m, a, nmax):
    """The sub-cycle search of non-periodic rows with held end cells, m (n, R), a (n-1, R)
    interior faces scaled in place per retry (kernels.py:169-199 with periodic=False).
    Returns nloop (R,) or None."""
    rows = m.shape[1]
    nloop = np.ones(rows, dtype=np.int64)
    todo, n = np.arange(rows), 1
    while True:
        mx, f = m[:, todo].copy(), a[:, todo]
        passed = np.ones(todo.size, dtype=bool)
        for _ in range(n):
            c = np.abs(f / pick_upwind(f, mx[:-1], mx[1:]))
            passed &= np.all(c < CFL_ONE, axis=0)
            mx[1:-1] = mass_after(mx[1:-1], f)
        nloop[todo] = n
        todo = todo[~passed]
        if todo.size == 0:
            return nloop
        n += 1
        a[:, todo] = shrink_progressive(a[:, todo], n)
        if n > nmax:
            return None


def _child_row_substeps(m, a, q4, nloop, limits):
    """nloop sub-steps of non-periodic rows, ends held (advect.py:347-366)."""
    rm, rxm, rym, rzm = q4
    af = a[..., None]
    for _ in range(nloop):
        mnew = mass_after(m[1:-1], a)
        minv = (1.0 / m)[..., None]
        alpha = af * pick_upwind(af, minv[:-1], minv[1:])
        ups = [pick_upwind(af, q[:-1], q[1:]) for q in (rm, rxm, rym, rzm)]
        f, pf, fy, fz = face_fluxes(af, alpha, *ups)
        rm[1:-1] = rm[1:-1] - (f[1:] - f[:-1])
        rxm[1:-1] = slope_update(rxm[1:-1], pf, af, rm[1:-1], f, m[1:-1, ..., None], mnew[..., None])
        if limits:
            rxm[1:-1] = limit(rxm[1:-1], rm[1:-1])
        rym[1:-1] = rym[1:-1] - (fy[1:] - fy[:-1])
        rzm[1:-1] = rzm[1:-1] - (fz[1:] - fz[:-1])
        m[1:-1] = mnew


def x_child(tree, st, r, js, je, ls, le, limits, max_nloop=50):
    """dynamu of a child: compress_xedges, then every (j, l) row's cells xr-1..im-xr+2 with the
    ends held and its own nloop (advect.py:301-344)."""
    compress_xedges(tree, st, r, js, je, ls, le)
    g, f = tree.g(r), st[r]
    xr = tree.ref(r, "x")
    i0, i1 = xr, g.im - xr + 1
    q4 = [f[k] for k in ("rm", "rxm", "rym", "rzm")]
    if limits:
        q4[1][i0 - 1 : i1 + 1, js:je, ls:le] = limit(q4[1][i0 - 1 : i1 + 1, js:je, ls:le], q4[0][i0 - 1 : i1 + 1, js:je, ls:le])
    nrow_l = le - ls + 1
    for j in range(js, je + 1):
        m = f["m"][i0 - 1 : i1 + 1, j, ls:le]
        a = f["am"][i0 - 1 : i1, j, ls:le]
        rows = [q[i0 - 1 : i1 + 1, j, ls:le] for q in q4]
        nloop = _row_search(m, a, max_nloop - 1)
        if nloop is None:
            raise CFLError("dynuamu: nloop too high")
        for n in np.unique(nloop):
            sel = np.flatnonzero(nloop == n)
            if sel.size == nrow_l:
                _child_row_substeps(m, a, rows, int(n), limits)
                continue
            parts = [q[:, sel].copy() for q in (m, *rows)]
            _child_row_substeps(parts[0], a[:, sel], parts[1:], int(n), limits)
            for q, part in zip((m, *rows), parts):
                q[:, sel] = part
        a[...] = a * nloop


def y_child(tree, st, r, is_, ie, ls, le, limits, max_nloop=10):
    """dynamv of a child: compress_yedges, then per layer the columns is..ie with cells
    yr-1..jm-yr+2 (ends held), one nloop per layer found with the whole layer of bm scaled
    (advect.py:369-432)."""
    compress_yedges(tree, st, r, is_, ie, ls, le)
    g, f = tree.g(r), st[r]
    yr = tree.ref(r, "y")
    j0, j1 = yr, g.jm - yr + 1
    rm, rxm, rym, rzm = (f[k] for k in ("rm", "rxm", "rym", "rzm"))
    if limits:
        rym[is_:ie, j0 - 1 : j1 + 1, ls:le] = limit(rym[is_:ie, j0 - 1 : j1 + 1, ls:le], rm[is_:ie, j0 - 1 : j1 + 1, ls:le])
    for lev in range(ls, le + 1):
        cells = f["m"][is_:ie, j0 - 1 : j1 + 1, lev]  # (ni, n)
        b = f["bm"][is_:ie, j0 : j1 + 1, lev]  # (ni, n-1)
        nloop = 1
        while True:
            mx, ok = cells.copy(), True
            for _ in range(nloop):
                beta = np.where(b >= 0.0, b / mx[:, :-1], b / mx[:, 1:])
                if np.any(np.abs(beta) >= CFL_ONE):
                    ok = False
                    continue
                mx[:, 1:-1] = (mx[:, 1:-1] + b[:, :-1]) - b[:, 1:]
            if ok:
                break
            layer = f["bm"][:, :, lev]
            layer[...] = shrink_progressive(layer, nloop + 1)
            nloop += 1
            if nloop == max_nloop:
                raise CFLError("nloop too high in dynamv")
        for _ in range(nloop):
            mnew = (cells[:, 1:-1] + b[:, :-1]) - b[:, 1:]
            inv_mnew = (1.0 / mnew)[..., None]
            flow_conv = (b[:, :-1] - b[:, 1:])[..., None]
            flow_mean = (3.0 * (b[:, :-1] + b[:, 1:]))[..., None]
            mass_x3 = (3.0 * cells[:, 1:-1])[..., None]
            bf = b[..., None]
            beta = bf / np.where(bf >= 0.0, cells[:, :-1, None], cells[:, 1:, None])
            q = [x[is_:ie, j0 - 1 : j1 + 1, lev] for x in (rm, rym, rxm, rzm)]
            ups = [np.where(bf >= 0.0, x[:, :-1], x[:, 1:]) for x in q]
            fl, pf, fx, fz = face_fluxes(bf, beta, *ups)
            rmq, rymq, rxmq, rzmq = (x[:, 1:-1] for x in q)
            rm_new = rmq + (fl[:, :-1] - fl[:, 1:])
            rmq[...] = rm_new
            f_sum = fl[:, :-1] + fl[:, 1:]
            rym_new = rymq + ((((pf[:, :-1] - pf[:, 1:]) - flow_conv * rymq) + flow_mean * rm_new) - mass_x3 * f_sum) * inv_mnew
            rymq[...] = limit(rym_new, rm_new) if limits else rym_new
            rxmq[...] = rxmq + (fx[:, :-1] - fx[:, 1:])
            rzmq[...] = rzmq + (fz[:, :-1] - fz[:, 1:])
            cells[:, 1:-1] = mnew
        layer = f["bm"][:, :, lev]
        layer[...] = layer * nloop


def column_trial(full_m, full_c, si, sj, iter_limit):
    """dynamw's search over every column of the region (im, jm, lm), c = cm / n fresh, the trial
    mass advanced on the scope (si, sj) only (advect.py:415-441).  Returns (n_advz or None,
    the trial mass)."""
    n = 1
    while True:
        c = full_c / float(n)
        mx, ok = full_m.copy(), True
        for _ in range(n):
            ci = c[..., 1:-1]
            gamma = np.abs(np.where(ci >= 0.0, ci / mx[..., :-1], ci / mx[..., 1:]))
            if np.any(gamma >= CFL_ONE):
                ok = False
                break
            mx[si, sj] = np.moveaxis(mass_after(np.moveaxis(mx[si, sj], 2, 0), np.moveaxis(c[si, sj], 2, 0)), 0, 2)
        if ok:
            return n, mx
        n += 1
        if n > iter_limit:
            return None, mx


def _z_substeps(m, rm, rxm, rym, rzm, cm, n_advz, limits):
    from .advection import z_substeps

    z_substeps(m, rm, rxm, rym, rzm, cm, n_advz, limits)


def z_child(tree, st, r, is_, ie, js, je, limits, iter_limit=15):
    """dynamw of a child: the search over the whole region, the sub-steps on the scope
    (advect.py:444-472)."""
    g, f, lm = tree.g(r), st[r], tree.lm
    im, jm = g.im, g.jm
    cm = f["cm"]
    if np.any(cm[:, :, 0] != 0.0) or np.any(cm[:, :, lm] != 0.0):
        raise ValueError("dynamw: found non-zero halo interfaces in cm")
    rm, rzm = f["rm"], f["rzm"]
    if limits:
        rzm[is_:ie, js:je, 1:lm] = limit(rzm[is_:ie, js:je, 1:lm], rm[is_:ie, js:je, 1:lm])
    n_advz, _ = column_trial(f["m"][1:im, 1:jm, 1:lm], cm[1:im, 1:jm, 0:lm], slice(is_ - 1, ie), slice(js - 1, je), iter_limit)
    if n_advz is None:
        raise CFLError("dynamw: too many CFL violations!")
    cells = (slice(is_, ie), slice(js, je), slice(1, lm))
    faces = (slice(is_, ie), slice(js, je), slice(0, lm))
    views = [f[k][cells] for k in ("m", "rm", "rxm", "rym", "rzm")]
    _z_substeps(*views, cm[faces], n_advz, limits)


# ------------------------------------------------------------------ physics (zoom/physics.py)
def scope(tree, r):
    """The interior scope as 0-based slices of an (im, jm) array (physics.py:19-21)."""
    return slice(tree.isr[r] - 1, tree.ier[r]), slice(tree.jsr[r] - 1, tree.jer[r])


def owned(tree, r):
    """(im, jm) bool: the scope's cells the region owns (physics.py:24-31)."""
    g = tree.g(r)
    own = np.zeros((g.im, g.jm), dtype=bool)
    own[scope(tree, r)] = True
    return own & (tree.zoomed[r] == r)


def convec_region(tree, st, r, conv, ndyn):
    """convec on the owned columns, dt = ndyn / (2 tref) (physics.py:34-52)."""
    g, f, lm = tree.g(r), st[r], tree.lm
    cols = np.nonzero(owned(tree, r))
    views = [f[k][1 : g.im, 1 : g.jm, 1:lm] for k in ("m", "rm", "rxm", "rym", "rzm")]
    work = [v[cols][:, None].copy() for v in views]
    met = {k: np.asarray(conv[k])[cols][:, None] for k in ("entu", "detu", "entd", "detd", "dkg", "cloud_top", "cloud_lfs")}
    convec(*work, met["entu"], met["detu"], met["entd"], met["detd"], met["cloud_top"], met["cloud_lfs"], met["dkg"],
           ndyn / (2.0 * g.tref), conv["lmax"])  # fmt: skip
    for v, w in zip(views, work):
        v[cols] = w[:, 0]


def sources_region(tree, st, r, injections):
    """rm += x, rzm -= x on the owned cells for each (x, level, tracer) (physics.py:55-64)."""
    g, f = tree.g(r), st[r]
    own = owned(tree, r)
    for x, level, t in injections:
        rm = f["rm"][1 : g.im, 1 : g.jm, level, t]
        rzm = f["rzm"][1 : g.im, 1 : g.jm, level, t]
        rm[own] = rm[own] + x[own]
        rzm[own] = rzm[own] - x[own]


def chemistry_region(tree, st, r, decay, tracers):
    """rm..rzm of the tracers times the scope's decay on the owned cells, 1 elsewhere
    (physics.py:67-78)."""
    g, f = tree.g(r), st[r]
    factor = np.ones((g.im, g.jm, tree.lm))
    factor[scope(tree, r)] = decay
    factor[~owned(tree, r)] = 1.0
    for t in tracers:
        for k in ("rm", "rxm", "rym", "rzm"):
            q = f[k][1 : g.im, 1 : g.jm, 1 : tree.lm, t]
            q[...] = q * factor


# ------------------------------------------------------------------ the driver (zoom/steps.py)
def put_edges(tree, st, r):
    put_xedges(tree, st, r)
    put_yedges(tree, st, r)


def do_steps(tree, st, r, conv, ndyn, limits=True, sources=None, chemistry=None):
    """The next operators of the region's split order up to a block boundary (where the edges
    are put unless the last was x) or a z right after y (steps.py:31-69)."""
    order = tree.order[r]
    for _ in range(NOPS):
        op = order[tree.status[r]]
        if op == "x":
            advect_x_zoom(tree, st, r, limits)
        elif op == "y":
            advect_y_zoom(tree, st, r, limits)
        elif op == "z":
            advect_z_zoom(tree, st, r, limits)
        elif op == "v":
            convec_region(tree, st, r, conv[r], ndyn)
        elif op == "s":
            if sources is not None:
                sources(tree, st, r)
        elif op == "c":
            if chemistry is not None:
                chemistry(tree, st, r)
        else:
            raise ValueError(f"do_steps: strange value in splitorderzoom: {op!r}")
        tree.status[r] += 1
        s = tree.status[r]
        if s % NOPS == 0:
            if op != "x":
                put_edges(tree, st, r)
            return
        if op == "z" and order[s - 2] == "y":
            return


def proces_region(tree, st, r, conv, ndyn, **kw):
    """tref sub-steps of do_steps, each followed by the children; then update_parent and
    m2phlb (m2phlb1 for region 1) (steps.py:72-82)."""
    for _ in range(tree.ref(r, "t")):
        do_steps(tree, st, r, conv, ndyn, **kw)
        for c in tree.children[r]:
            proces_region(tree, st, c, conv, ndyn, **kw)
        if tree.status[r] % NOPS != 0:
            do_steps(tree, st, r, conv, ndyn, **kw)
    if r != 1:
        update_parent(tree, st, r)
        m2phlb(tree, st, r)
    else:
        m2phlb1(tree, st, 1)


def step(tree, st, conv, ndyn, **kw):
    """One ndyn step: statuses reset, Proces_Region(1) for each half (steps.py:85-88)."""
    tree.status = [0] * (tree.n + 1)
    for _ in range(2):
        proces_region(tree, st, 1, conv, ndyn, **kw)


# ------------------------------------------------------------------ Check_CFL (zoom/check_cfl.py)
class CFLRun:
    """regionm_status, cfl_ok and the CFL split order of one Check_CFL (check_cfl.py:63-74)."""

    def __init__(self, tree, caps):
        self.caps = caps
        self.order = tree._orders(CFL_SPLITORDER, CFL_NOPS)
        self.status = [0] * (tree.n + 1)
        self.ok = [True] * (tree.n + 1)

    def all_ok(self):
        return all(self.ok[1:])


def _m_views(tree, st, r):
    g, f, lm = tree.g(r), st[r], tree.lm
    return f["m"][1 : g.im, 1 : g.jm, 1:lm]


def mass_x_zoom(tree, st, run, r, cap):
    """advectmxzoom (check_cfl.py:85-98)."""
    put_xedges(tree, st, r, tracers=False)
    blk, second = block(tree, r, run.status[r], CFL_NOPS, run.order[r])
    js, je = x_scope(tree, r, blk)
    saved = _outer_x(tree, st, r, second)
    if r == 1:
        ok = mass_x_region1(tree, st, js, je, cap)
    else:
        ok = mass_x_child(tree, st, r, js, je, 1, tree.lm, cap)
    run.ok[r] = ok
    _restore_outer_x(tree, st, r, saved)


def mass_x_region1(tree, st, js, je, cap):
    """dynamum(1) (on the reduced grid when there is one, pydyn/cfl.py:30-57) and the halos
    m(0) = m(im) of rows js..je, m(im+1) = m(1) on the uniform rows (check_cfl.py:101-118)."""
    g, f, lm = tree.g(1), st[1], tree.lm
    im, jm = g.im, g.jm
    m_int, am_int = f["m"][1:im, 1:jm, 1:lm], f["am"][0:im, 1:jm, 1:lm]
    grid = g.redgrid
    if grid is None:
        if not mass_x(m_int, am_int, cap):
            return False
    else:
        m_uni, am_uni = m_int.copy(), am_int.copy()
        uni2red_mass(grid, m_int)
        uni2red_mf(grid, am_int)
        for j in range(jm):
            n = grid.imred[j]
            if not mass_x(m_int[:n, j : j + 1], am_int[: n + 1, j : j + 1], cap):
                return False
        red2uni_mass(grid, m_int, (m_uni + am_uni[:-1]) - am_uni[1:])
        am_int[...] = am_uni
    reduced = set(grid.rows) if grid is not None else set()
    m = f["m"]
    for j in range(js, je + 1):
        m[0, j] = m[im, j].copy()
        if j - 1 not in reduced:
            m[im + 1, j] = m[1, j].copy()
    return True


def mass_x_child(tree, st, r, js, je, ls, le, cap):
    """dynamum of a child, mass only (check_cfl.py:121-145): compress, per row search and
    update, uncompress and mix -- not reached when a row fails."""
    compress_xedges(tree, st, r, js, je, ls, le, tracers=False)
    g, f, xr = tree.g(r), st[r], tree.ref(r, "x")
    i0, i1 = xr, g.im - xr + 1
    for j in range(js, je + 1):
        cells, a = f["m"][i0 - 1 : i1 + 1, j, ls:le], f["am"][i0 - 1 : i1, j, ls:le]
        nloop = _row_search(cells, a, cap - 1)
        if nloop is None:
            return False
        for v in np.unique(nloop):
            sel = np.flatnonzero(nloop == v)
            part, faces = cells[:, sel].copy(), a[:, sel]
            for _ in range(int(v)):
                part[1:-1] = mass_after(part[1:-1], faces)
            cells[:, sel] = part
        a[...] = a * nloop
    uncompress_xedges(tree, st, r, js, je, ls, le, tracers=False)
    mix_edges(tree, st, r, tracers=False)
    return True


def mass_y_zoom(tree, st, run, r, cap):
    """advectmyzoom (check_cfl.py:148-161)."""
    put_yedges(tree, st, r, tracers=False)
    blk, second = block(tree, r, run.status[r], CFL_NOPS, run.order[r])
    is_, ie = y_scope(tree, r, blk)
    saved = _outer_y(tree, st, r, second)
    if r == 1:
        g, f, lm = tree.g(1), st[1], tree.lm
        ok = mass_y(f["m"][1 : g.im, 1 : g.jm, 1:lm], f["bm"][1 : g.im, 1 : g.jm + 1, 1:lm], cap)
    else:
        ok = mass_y_child(tree, st, r, is_, ie, 1, tree.lm, cap)
    run.ok[r] = ok
    _restore_outer_y(tree, st, r, saved)


def mass_y_child(tree, st, r, is_, ie, ls, le, cap):
    """dynamvm of a child (check_cfl.py:171-201): on a failed trial only the scope's faces are
    scaled, but the whole layer is multiplied back."""
    compress_yedges(tree, st, r, is_, ie, ls, le, tracers=False)
    g, f, yr = tree.g(r), st[r], tree.ref(r, "y")
    j0, j1 = yr, g.jm - yr + 1
    for lev in range(ls, le + 1):
        cells, b = f["m"][is_:ie, j0 - 1 : j1 + 1, lev], f["bm"][is_:ie, j0 : j1 + 1, lev]
        nloop = 1
        while True:
            mx, ok = cells.copy(), True
            for _ in range(nloop):
                beta = np.abs(np.where(b >= 0.0, b / mx[:, :-1], b / mx[:, 1:]))
                if (beta >= CFL_ONE).any():
                    ok = False
                    break
                mx[:, 1:-1] = (mx[:, 1:-1] + b[:, :-1]) - b[:, 1:]
            if ok:
                break
            b[...] = shrink_progressive(b, nloop + 1)
            nloop += 1
            if nloop > cap:
                return False
        for _ in range(nloop):
            cells[:, 1:-1] = (cells[:, 1:-1] + b[:, :-1]) - b[:, 1:]
        layer = f["bm"][:, :, lev]
        layer[...] = layer * nloop
    uncompress_yedges(tree, st, r, is_, ie, ls, le, tracers=False)
    mix_edges(tree, st, r, tracers=False)
    return True


def mass_z_zoom(tree, st, run, r, cap):
    """advectmzzoom / dynamwm (check_cfl.py:204-226)."""
    blk, _ = block(tree, r, run.status[r], CFL_NOPS, run.order[r])
    is_, ie, js, je = z_scope(tree, r, blk)
    g, f, lm = tree.g(r), st[r], tree.lm
    full = f["m"][1 : g.im, 1 : g.jm, 1:lm]
    si, sj = slice(is_ - 1, ie), slice(js - 1, je)
    n, mx = column_trial(full, f["cm"][1 : g.im, 1 : g.jm, 0:lm], si, sj, cap)
    run.ok[r] = n is not None
    if n is not None:
        full[si, sj] = mx[si, sj]
        if r != 1:
            mix_edges(tree, st, r, tracers=False)


def cfl_do_steps(tree, st, run, r):
    """The next operators up to a block boundary of three, back at the first failure
    (check_cfl.py:232-245)."""
    ops = {"x": mass_x_zoom, "y": mass_y_zoom, "z": mass_z_zoom}
    for _ in range(CFL_NOPS):
        op = run.order[r][run.status[r]]
        ops[op](tree, st, run, r, run.caps["xyz".index(op)])
        if not run.all_ok():
            return
        run.status[r] += 1
        if run.status[r] % CFL_NOPS == 0:
            return


def determine_cfl_iter(tree, st, run, r):
    """The region's tref blocks each followed by its children, then updatem_parent; any failure
    unwinds straight back (check_cfl.py:248-266)."""
    run.ok[r] = True
    for _ in range(tree.ref(r, "t")):
        cfl_do_steps(tree, st, run, r)
        if not run.ok[r]:
            return
        for c in tree.children[r]:
            if not run.all_ok():
                break
            determine_cfl_iter(tree, st, run, c)
        if run.status[r] % CFL_NOPS != 0 and run.all_ok():
            cfl_do_steps(tree, st, run, r)
            if not run.ok[r]:
                return
    if r != 1:
        update_parent(tree, st, r, tracers=False)


def check_cfl_zoom(tree, st, ndyn, n, nread, interpolate=None, max_iteration=100, caps=CAPS):
    """Check_CFL for the tree (check_cfl.py:269-290): per step Setup_MassFlow and the xyz / zyx
    rehearsals; m, am, bm, cm restored; on a failure the next valid ndyn and the fluxes scaled.
    Returns (ndyn, n, iterations)."""
    tref_max = max(tree.g(r).tref for r in tree.regions())
    for iteration in range(1, max_iteration + 1):
        run = CFLRun(tree, caps)
        saved = {(r, k): st[r][k].data.copy() for r in tree.regions() for k in ("m", "am", "bm", "cm")}
        for i in range(1, n + 1):
            if interpolate is not None:
                interpolate(tree, st, i, ndyn)
            setup_massflow(tree, st, ndyn)
            determine_cfl_iter(tree, st, run, 1)
            if run.all_ok():
                determine_cfl_iter(tree, st, run, 1)
            if not run.all_ok():
                break
            run.status = [0] * (tree.n + 1)
        for (r, k), v in saved.items():
            st[r][k].data[...] = v
        if run.all_ok():
            return ndyn, n, iteration
        ndyn_old = ndyn
        ndyn = new_valid_timestep(ndyn, nread, nread, tref_max)
        n = nint(float(n * ndyn_old) / float(ndyn))
        fraction = float(ndyn) / float(ndyn_old)
        for f in st.values():
            for k in ("am", "bm", "cm"):
                f[k].data[...] = f[k].data * fraction
    raise RuntimeError("check_cfl_zoom: exceeded maximum number of time steps")


# ------------------------------------------------------------------ a run of the tree
SEC_MONTH = 30.0 * 86400.0
XMCH4 = 12.01115 + 1.0079 * 4.0


def tree_of(parent, child, lm, at, bt, redgrid=None):
    """The Tree of a global parent grid (its x advection on the reduced grid (jred, clustsize)
    when given) and one pydyn_amd.meteo.Child."""
    c = child
    grid = None if redgrid is None else RedGrid(parent.im, parent.jm, *redgrid)
    glb = Region("globe", parent.im, parent.jm, xcyc=1, touch_sp=1, touch_np=1, redgrid=grid)
    kid = Region(c.name, c.im, c.jm, c.xbeg, c.xend, c.ybeg, c.yend, c.xref, c.yref, 1, c.tref, parent=1)
    tree = Tree([glb, kid], lm, at, bt)
    if (tree.ibeg[2], tree.jbeg[2], tree.iend[2], tree.jend[2]) != (c.ibeg, c.jbeg, c.iend, c.jend):
        raise ValueError("the child's declared bounds do not snap to its footprint")
    return tree


class ZoomRun:
    """A global region with one child on nested meteo (pydyn_amd.meteo.zoom): Check_CFL for the
    tree, then the steps, with sources (level 1 of tracer 1, month2dt of the region's half
    step) and chemistry (the region's decay over its scope, dt = ndyn / (2 tref)) as callbacks
    on every region's s and c (pydyn/forward.py:411-451 generalised to the tree)."""

    def __init__(self, pmet, cmet, limits=True, redgrid=None, ensemble=None):
        g = pmet.g
        self.mets = {1: pmet, 2: cmet}
        self.tree = tree_of(g, cmet.c, g.lm, g.at, g.bt, redgrid)
        self.limits = limits
        self.ensemble = ensemble
        self.nt = pmet.nt if ensemble is None else ensemble.nt
        tree, lm = self.tree, g.lm
        self.st = new_state(tree, self.nt)
        for r, met in self.mets.items():
            f, gr = self.st[r], tree.g(r)
            init = met.state() if ensemble is None else ensemble.state(met)
            f["m"][1 : gr.im, 1 : gr.jm, 1:lm] = init["m"]
            for k in ("rm", "rxm", "rym", "rzm"):
                f[k][1 : gr.im, 1 : gr.jm, 1:lm] = init[k]
            f["dxyp"][1 : gr.jm] = met.dxyp if r > 1 else g.dxyp
        # the parent's footprint holds its child's air and tracer mass from the start
        update_parent(tree, self.st, 2)
        self.conv, self.emis, self.loss = {}, {}, {}
        for r, met in self.mets.items():
            gr = tree.g(r)
            meteo_m = met.m
            self.conv[r] = {
                "entu": met.entu, "detu": met.detu, "entd": met.entd, "detd": met.detd,
                "dkg": physics.dkg(met.kvh, meteo_m, met.gph, g.lmax),
                "cloud_top": met.cloud_top, "cloud_lfs": met.cloud_lfs, "lmax": g.lmax,
            }  # fmt: skip
            self.emis[r] = met.emis
            si, sj = scope(tree, r)
            self.loss[r] = met.loss[si, sj]
        self._set_fluxes(0.0)
        setup_massflow(tree, self.st, 3600)
        m2phlb(tree, self.st, 2)
        m2phlb1(tree, self.st, 1)
        self.t = 0.0
        self.log = []

    def _set_fluxes(self, t_sec):
        for r, met in self.mets.items():
            gr, f, lm = self.tree.g(r), self.st[r], self.tree.lm
            pu, pv = met.fluxes(t_sec)
            f["pu"][0 : gr.im, 1 : gr.jm, 1:lm] = pu
            f["pv"][1 : gr.im, 1 : gr.jm + 1, 1:lm] = pv

    def sources(self, ndyn):
        """Each region's emission at level 1: of tracer 1, or with an ensemble of every member's
        first tracer, the field scaled by the member's factor first."""

        def add(tree, st, r):
            factor = physics.month2dt(ndyn, SEC_MONTH, tree.g(r).tref)
            if self.ensemble is None:
                injections = [(physics.inject(self.emis[r], factor, XMCH4, XMCH4, 1.0), 1, 1)]
            else:
                injections = [(physics.inject(self.emis[r] * s, factor, XMCH4, XMCH4, 1.0), 1, t + 1)
                              for t, s in enumerate(self.ensemble.escale()) if s != 0.0]  # fmt: skip
            sources_region(tree, st, r, injections)

        return add

    def chemistry(self, ndyn):
        def apply(tree, st, r):
            factor = physics.decay(self.loss[r], ndyn / (2.0 * tree.g(r).tref))
            chemistry_region(tree, st, r, factor, range(1, self.nt + 1))

        return apply

    def interval(self, nread=3600, ndyn_max=3600):
        t0, tree = self.t, self.tree

        def interpolate(tree_, st, i, ndyn):
            self._set_fluxes(t0 + (i - 0.5) * ndyn)

        ndyn, n, _ = check_cfl_zoom(tree, self.st, ndyn_max, nread // ndyn_max, nread, interpolate)
        for s in range(n):
            self._set_fluxes(t0 + s * ndyn + ndyn / 2.0)
            setup_massflow(tree, self.st, ndyn)
            step(tree, self.st, self.conv, ndyn, limits=self.limits, sources=self.sources(ndyn),
                 chemistry=self.chemistry(ndyn))  # fmt: skip
        self.t = t0 + n * ndyn
        self.log.append((t0, ndyn, n))
        return ndyn, n
