"""The region tree of a zoom configuration (pydyn/zoom/tree.py): the regions as declared, and
what the model's initexit derives from them -- children, footprints, interior scopes, the
zoomed / edge maps and each region's split order.

Host logic only; the device bodies (pydyn_amd.zoom.lib) receive the derived numbers in each
region's descriptor.
"""

from dataclasses import dataclass

import numpy as np

__all__ = ["N_OPERATORS", "SPLITORDER", "RedGrid", "Region", "Zoom", "split_orders"]

SPLITORDER = "xyzvsccsvzyx"  # the model's operator order of a step (pydyn/zoom/tree.py:18)
N_OPERATORS = len(SPLITORDER) // 2


def _nint(x):
    """Fortran nint: half away from zero."""
    r = int(np.floor(abs(x) + 0.5))
    return r if x >= 0 else -r


class RedGrid:
    """A reduced grid for a region's x advection (pydyn/redgrid.py:26-54): on row jred[k]
    (1-based) clusters of clustsize[k] cells are merged, leaving imred = im / clustsize cells."""

    def __init__(self, im, jm, jred, clustsize):
        self.im, self.jm = im, jm
        self.jred = [int(j) for j in jred]
        self.imred = np.full(jm, im, dtype=np.int32)
        for j, c in zip(self.jred, clustsize):
            if im % c:
                raise ValueError(f"redgrid: im={im} not a multiple of clustsize {c}")
            self.imred[j - 1] = im // c

    @property
    def rows(self):
        """The reduced rows, 0-based."""
        return [j - 1 for j in self.jred]


@dataclass
class Region:
    """A region as declared (pydyn/zoom/tree.py:24-52): refinements relative to the globe,
    parent 0 = the globe."""

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
    redgrid: object = None  # RedGrid of the region's x advection (region 1 only), or None


def split_orders(zoom, base, nops):
    """Each region's operator order: its parent's blocks of `nops`, each repeated tref_ times
    with every second copy reversed (define_splitorderzoom, zoom_mode 1; pydyn/zoom/tree.py:
    162-176)."""
    out = {1: base}
    for r in range(2, zoom.nregions + 1):
        t = zoom.ref(r, "t")
        parent = out[zoom.g(r).parent]
        pieces = []
        for k in range(0, len(parent), nops):
            blk = parent[k : k + nops]
            pieces.extend(blk if c % 2 == 0 else blk[::-1] for c in range(t))
        out[r] = "".join(pieces)
    return out


class Zoom:
    """The tree (pydyn/zoom/tree.py:55-191).  Regions count from 1; status[r] is the position
    of region r in its split order during a step."""

    def __init__(self, regions, lm, at=None, bt=None, splitorder=SPLITORDER):
        self.region = [None, *regions]
        self.nregions = len(regions)
        self.lm = lm
        self.at = np.zeros(lm + 1) if at is None else np.asarray(at, dtype=float)
        self.bt = np.zeros(lm + 1) if bt is None else np.asarray(bt, dtype=float)
        self.status = [0] * (self.nregions + 1)
        self.children = {r: [] for r in range(self.nregions + 1)}
        for r in range(2, self.nregions + 1):
            self.children[regions[r - 1].parent].append(r)
        self._place_children()
        self._scopes()
        self.zoomed, self.edge = {}, {}
        for r in self.regions():
            self.zoomed[r], self.edge[r] = self._maps(r)
        for r in range(2, self.nregions + 1):
            t = self.ref(r, "t")
            if t > 4 or t == 3:
                raise ValueError(f"define_splitorderzoom: region {r}: tref_ = {t} not 1, 2 or 4")
        self.splitorderzoom = split_orders(self, splitorder, N_OPERATORS)
        self._supported()

    def regions(self):
        return range(1, self.nregions + 1)

    def g(self, r):
        return self.region[r] if r > 0 else Region("globe", self.region[1].im, self.region[1].jm)

    def ref(self, r, axis):
        """The refinement of r relative to its parent (integer division, as the model's)."""
        return getattr(self.g(r), axis + "ref") // getattr(self.g(self.g(r).parent), axis + "ref")

    def _place_children(self):
        """ibeg..iend x jbeg..jend: each child's footprint in its parent's cells, its declared
        bounds snapped by nint (determine_children_etc)."""
        g1 = self.g(1)
        dx1 = (g1.xend - g1.xbeg) / g1.im * g1.xref
        dy1 = (g1.yend - g1.ybeg) / g1.jm * g1.yref
        self.ibeg, self.iend, self.jbeg, self.jend, self.lbeg, self.lend = {}, {}, {}, {}, {}, {}
        for r in range(2, self.nregions + 1):
            g, p = self.g(r), self.g(self.g(r).parent)
            sx, sy = dx1 / p.xref, dy1 / p.yref
            self.ibeg[r] = _nint((g.xbeg - p.xbeg) / sx) + 1
            self.iend[r] = _nint((g.xend - p.xbeg) / sx)
            self.jbeg[r] = _nint((g.ybeg - p.ybeg) / sy) + 1
            self.jend[r] = _nint((g.yend - p.ybeg) / sy)
            self.lbeg[r], self.lend[r] = 1, self.lm
            ib, ie = self.ibeg[r], self.iend[r]
            ncell = ie - ib + 1 if ie > ib else p.im - ib + 1 + ie
            if ncell * self.ref(r, "x") != g.im:
                raise ValueError(f"region {r}: check xbeg/xend, im and ibeg/iend")
            if (self.jend[r] - self.jbeg[r] + 1) * self.ref(r, "y") != g.jm:
                raise ValueError(f"region {r}: check ybeg/yend, jm and jbeg/jend")

    def _scopes(self):
        """isr..ier x jsr..jer: the interior a region owns outright (the ring of parent-cell
        blocks along a child's walls excluded)."""
        g1 = self.g(1)
        self.isr, self.ier, self.jsr, self.jer = {1: 1}, {1: g1.im}, {1: 1}, {1: g1.jm}
        for r in range(2, self.nregions + 1):
            g, xr, yr = self.g(r), self.ref(r, "x"), self.ref(r, "y")
            self.isr[r], self.ier[r] = (1, g.im) if g.xcyc != 0 else (xr + 1, g.im - xr)
            self.jsr[r] = 1 if g.touch_sp == 1 else yr + 1
            self.jer[r] = g.jm if g.touch_np == 1 else g.jm - yr

    def _maps(self, r):
        """(zoomed, edge) of region r, (im, jm) int32, 0-based (blank_zoom_region): zoomed is
        the region that owns each cell (a child strictly inside its footprint); edge marks the
        cells on a child's walls (1 x walls, +2 y walls) and, in a child, its own ring (-1)."""
        g = self.g(r)
        zoomed = np.full((g.im, g.jm), r, dtype=np.int32)
        edge = np.zeros((g.im, g.jm), dtype=np.int32)
        ii = np.arange(1, g.im + 1)[:, None]
        jj = np.arange(1, g.jm + 1)[None, :]
        for c in self.children[r]:
            gc = self.g(c)
            ib, ie, jb, je = self.ibeg[c], self.iend[c], self.jbeg[c], self.jend[c]
            in_i = (gc.xcyc == 1) | ((ii > ib) & (ii < ie))
            in_j = ((gc.touch_sp == 1) | (jj > jb)) & ((gc.touch_np == 1) | (jj < je))
            zoomed[np.broadcast_to(in_i & in_j, zoomed.shape)] = c
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

    def _supported(self):
        """The configurations the model runs (pydyn/zoom/tree.py:178-191): a global region 1,
        children strictly inside their parent, neither cyclic nor touching a pole."""
        g1 = self.g(1)
        if not (g1.xcyc == 1 and g1.touch_sp == 1 and g1.touch_np == 1):
            raise NotImplementedError("region 1 must be global: cyclic and touching both poles")
        for r in range(2, self.nregions + 1):
            g, p = self.g(r), self.g(self.g(r).parent)
            if g.xcyc or g.touch_sp or g.touch_np:
                raise NotImplementedError(f"region {r}: cyclic or pole-touching children are not supported")
            if self.ibeg[r] >= self.iend[r]:
                raise NotImplementedError(f"region {r}: a child across the dateline is not supported")
            if self.ibeg[r] <= 1 or self.iend[r] >= p.im:
                raise NotImplementedError(f"region {r}: put_xedges would read the parent's halo")
            if self.jbeg[r] <= 1 or self.jend[r] >= p.jm:
                raise NotImplementedError(f"region {r}: put_yedges would read the parent's halo")
            if g.redgrid is not None:
                raise NotImplementedError(f"region {r}: a reduced grid in a child is not supported")

    def scope(self, r):
        """The interior scope as 0-based slices of an (im, jm) array."""
        return slice(self.isr[r] - 1, self.ier[r]), slice(self.jsr[r] - 1, self.jer[r])

    def owned(self, r):
        """(im, jm) bool: the cells of the scope that region r owns (zoomed == r)."""
        g = self.g(r)
        own = np.zeros((g.im, g.jm), dtype=bool)
        own[self.scope(r)] = True
        return own & (self.zoomed[r] == r)
