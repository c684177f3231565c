"""Every region's arrays on the device, with the Fortran bounds the model declares
(pydyn/zoom/state.py, pydyn/layout.py:26-36).

A Field is a float64 device tensor of the declared logical shape, i fastest in memory, with
its lower bounds; `sec` takes inclusive Fortran sections as torch views.  `ZoomState.desc(r)`
is region r's tm5_zregion for the library (built once: the arrays never move).
"""

import ctypes

import numpy as np

from .. import _lib
from .. import device as D

__all__ = ["FIELDS", "Field", "HostState", "ZoomState", "declared"]

FIELDS = ("m", "rm", "rxm", "rym", "rzm", "am", "bm", "cm", "pu", "pv", "sp", "phlb", "dxyp")


def declared(name, im, jm, lm, nt):
    """((lower, upper), ...) of a region's array (pydyn/layout.py:26-36)."""
    cells = ((-1, im + 2), (-1, jm + 2), (1, lm))
    faces = ((0, im + 1), (0, jm + 1), (0, lm + 1))
    return {
        "m": cells,
        "rm": (*cells, (1, nt)),
        "rxm": (*cells, (1, nt)),
        "rym": (*cells, (1, nt)),
        "rzm": (*cells, (1, nt)),
        "am": faces,
        "bm": faces,
        "cm": faces,
        "pu": ((0, im), (1, jm), (1, lm)),
        "pv": ((1, im), (1, jm + 1), (1, lm)),
        "sp": ((-1, im + 2), (-1, jm + 2), (1, 1)),
        "phlb": ((-1, im + 2), (-1, jm + 2), (1, lm + 1)),
        "dxyp": ((1, jm),),
    }[name]


class Field:
    """A device array addressed by inclusive Fortran sections."""

    def __init__(self, bounds):
        self.lower = tuple(lo for lo, _ in bounds)
        self.t = D.zeros(tuple(hi - lo + 1 for lo, hi in bounds))

    def sec(self, *key):
        """The torch view of an inclusive Fortran section; ints index, (lo, hi) pairs slice,
        axes left out are whole."""
        ix = []
        for lo, k in zip(self.lower, key):
            ix.append(slice(k[0] - lo, k[1] - lo + 1) if isinstance(k, tuple) else k - lo)
        return self.t[tuple(ix)]

    def hview(self):
        return D.hview(self.t, self.lower)

    def numel(self):
        return self.t.numel()

    def host(self):
        """The whole array in C order (the layout a NumPy state holds it in)."""
        return D.download(self.t)


class ZoomState:
    """Every region's Fields (`fields[r][name]`), nt tracers, and the device data the bodies
    need beyond the state: each region's owned and edge maps and scratch for saved arrays."""

    def __init__(self, zoom, nt):
        self.zoom, self.nt = zoom, nt
        torch = __import__("torch")
        lm = zoom.lm
        self.fields, self._desc, self._owned, self._edge, self.saved = {}, {}, {}, {}, {}
        self.slabs = {}
        for r in zoom.regions():
            g = zoom.g(r)
            self.fields[r] = {k: Field(declared(k, g.im, g.jm, lm, nt)) for k in FIELDS}
            self._owned[r] = D.upload(zoom.owned(r).astype(np.int32)).to(torch.int32)
            self._edge[r] = D.upload(zoom.edge[r].astype(np.int32)).to(torch.int32)
            n = 2 * (max(g.im, g.jm) + 2) * (lm + 2)
            self.slabs[r] = (D.empty((n,)), D.empty((n,)))  # outer x faces, outer y faces
        self.at = D.upload(np.asarray(zoom.at, dtype=np.float64))
        self.bt = D.upload(np.asarray(zoom.bt, dtype=np.float64))

    def __getitem__(self, r):
        return self.fields[r]

    def desc(self, r):
        """tm5_zregion of region r."""
        if r not in self._desc:
            zoom, f, g = self.zoom, self.fields[r], self.zoom.g(r)
            d = _lib.ZRegion()
            d.im, d.jm, d.lm, d.nt = g.im, g.jm, zoom.lm, self.nt
            d.xref, d.yref, d.zref, d.tref = (zoom.ref(r, a) for a in "xyzt")
            d.tref_abs = g.tref
            if r > 1:
                d.ibeg, d.iend, d.jbeg, d.jend = zoom.ibeg[r], zoom.iend[r], zoom.jbeg[r], zoom.jend[r]
            d.isr, d.ier, d.jsr, d.jer = zoom.isr[r], zoom.ier[r], zoom.jsr[r], zoom.jer[r]
            d.xcyc, d.touch_sp, d.touch_np = g.xcyc, g.touch_sp, g.touch_np
            d.whole = int(g.im // g.xref == zoom.g(1).im)
            for k in ("m", "rm", "rxm", "rym", "rzm", "am", "bm", "cm", "pu", "pv", "sp", "phlb"):
                setattr(d, k, f[k].hview())
            d.dxyp = f["dxyp"].t.data_ptr()
            d.owned = self._owned[r].data_ptr()
            d.edge = self._edge[r].data_ptr()
            grid = g.redgrid
            if grid is not None:
                torch = __import__("torch")
                jred = torch.tensor(grid.jred, dtype=torch.int32, device="cuda")
                imred = torch.tensor(grid.imred, dtype=torch.int32, device="cuda")
                scratch = D.empty((f["m"].numel() + f["am"].numel(),))
                self._keep = getattr(self, "_keep", []) + [jred, imred, scratch]
                d.nred, d.jred, d.imred, d.scratch = len(grid.jred), jred.data_ptr(), imred.data_ptr(), scratch.data_ptr()
            self._desc[r] = d
        return self._desc[r]

    def ptr(self, r):
        return ctypes.byref(self.desc(r))

    def upload_interior(self, r, init, dxyp):
        """m, rm..rzm of region r's interior from a NumPy state (pydyn layout), and dxyp."""
        g, lm = self.zoom.g(r), self.zoom.lm
        f = self.fields[r]
        for k in ("m", "rm", "rxm", "rym", "rzm"):
            f[k].sec((1, g.im), (1, g.jm), (1, lm)).copy_(D.upload(init[k]))
        f["dxyp"].t.copy_(D.upload(np.asarray(dxyp, dtype=np.float64)))

    def host(self):
        """{r: {name: C-ordered NumPy array with the declared shape}}."""
        return {r: {k: v.host() for k, v in f.items()} for r, f in self.fields.items()}


class HostState:
    """Every region's arrays on the host: `fields[r][name]` a pydyn_amd.dump.FArray (the declared
    shape in C order, its Fortran lower bounds), as the reference's host ZoomState holds FViews
    (pydyn/zoom/state.py:16-45); `zoom` the tree they belong to."""

    def __init__(self, fields, zoom=None):
        self.fields, self.zoom = fields, zoom

    def __getitem__(self, r):
        return self.fields[r]

    @property
    def nt(self):
        return self.fields[1]["rm"].data.shape[3]
