"""A zoom tree resident on the device, run interval by interval: a global region with one child
on nested synthetic meteo (pydyn_amd.meteo.zoom), Check_CFL of the tree, then the steps with
Setup_MassFlow, every region's convection, sources (its emission at level 1 of tracer 1,
month2dt of its half step ndyn / (2 tref)) and chemistry (the decay of its loss rates over
ndyn / (2 tref)) on the cells it owns.

Two drivers issue the same operations: the library's native one (tm5_tree_*, the default; it
also queues intervals as checked steps replayed from HIP graphs, ZoomRun.run) and the Python
restatement of the reference's driver over the body table (pydyn_amd.zoom.driver).
"""

import ctypes

import numpy as np

from .. import _lib
from .. import device as D
from ..gpu import kernels as K
from ..meteo import MOD_AMPLITUDE, MOD_PERIOD, SEC_MONTH, XMCH4
from . import driver as Z
from . import lib as L
from .state import ZoomState
from .tree import RedGrid, Region, Zoom

__all__ = ["NativeTree", "Physics", "ZoomRun", "tree_of"]


def tree_of(parent_grid, child, redgrid=None):
    """The Zoom of a global grid (its x advection on the reduced grid (jred, clustsize) when
    given) and one pydyn_amd.meteo.Child."""
    g, c = parent_grid, child
    grid = None if redgrid is None else RedGrid(g.im, g.jm, *redgrid)
    glb = Region("globe", g.im, g.jm, xcyc=1, touch_sp=1, touch_np=1, redgrid=grid)
    kid = Region(c.name, c.im, c.jm, c.xbeg, c.xend, c.ybeg, c.yend, c.xref, c.yref, 1, c.tref, parent=1)
    zoom = Zoom([glb, kid], g.lm, g.at, g.bt)
    if (zoom.ibeg[2], zoom.jbeg[2], zoom.iend[2], zoom.jend[2]) != (c.ibeg, c.jbeg, c.iend, c.jend):
        raise ValueError("the child's declared bounds do not snap to its footprint")
    return zoom


class Physics:
    """Every region's column-physics inputs on the device, and their tm5_zphysics per ndyn."""

    def __init__(self, zoom, st, mets, lmax, escale=None):
        torch = __import__("torch")
        self.zoom, self.st, self.lmax = zoom, st, lmax
        self.arrays, self._structs = {}, {}
        self.escale = None if escale is None else D.upload(np.asarray(escale, dtype=np.float64))
        for r, met in mets.items():
            up = D.upload
            a = {k: up(getattr(met, k)) for k in ("entu", "detu", "entd", "detd", "emis", "loss")}
            a["dkg"] = K.dkg(up(met.kvh), up(met.m), up(met.gph), lmax)
            a["cloud_top"] = up(met.cloud_top.astype(np.int32)).to(torch.int32)
            a["cloud_lfs"] = up(met.cloud_lfs.astype(np.int32)).to(torch.int32)
            self.arrays[r] = a

    def struct(self, r, ndyn):
        key = (r, ndyn)
        if key not in self._structs:
            a, tref = self.arrays[r], self.zoom.g(r).tref
            p = _lib.ZPhysics()
            for k in ("entu", "detu", "entd", "detd", "dkg", "emis", "loss"):
                setattr(p, k, D.view(a[k]))
            p.cloud_top, p.cloud_lfs = D.iview(a["cloud_top"]), D.iview(a["cloud_lfs"])
            p.lmax = self.lmax
            half = float(ndyn) / (2.0 * tref)
            p.dt_conv = half
            p.factor = half / SEC_MONTH  # month2dt (pydyn/emis.py:65-68)
            p.ratio, p.efrac = XMCH4 / XMCH4, 1.0
            p.level, p.tracer = 1, 1
            p.escale = None if self.escale is None else self.escale.data_ptr()
            p.loss_stride = a["loss"].stride(3)
            p.dt_chem = half
            p.at, p.bt = self.st.at.data_ptr(), self.st.bt.data_ptr()
            self._structs[key] = p
        return self._structs[key]


class NativeTree:
    """The library's driver of a ZoomRun's tree (tm5_tree): the regions' descriptors, their
    physics and steady fluxes handed over once; the state stays the ZoomRun's."""

    def __init__(self, run):
        zoom, st = run.zoom, run.st
        regions = list(zoom.regions())
        n = len(regions)
        self._regions = (_lib.ZRegion * n)(*(st.desc(r) for r in regions))
        self._parent = (ctypes.c_int * n)(*(zoom.g(r).parent if r > 1 else 0 for r in regions))
        self._physics = (_lib.ZPhysics * n)(*(run.phys.struct(r, 3600) for r in regions))
        self._pu0 = (_lib.View * n)(*(D.view(run.flux0[r][0]) for r in regions))
        self._pv0 = (_lib.View * n)(*(D.view(run.flux0[r][1]) for r in regions))
        d = _lib.TreeDesc()
        d.nregions = n
        d.regions, d.parent, d.physics = self._regions, self._parent, self._physics
        d.pu0, d.pv0 = self._pu0, self._pv0
        d.at, d.bt = st.at.data_ptr(), st.bt.data_ptr()
        d.mod_period, d.mod_amplitude, d.sec_month = MOD_PERIOD, MOD_AMPLITUDE, SEC_MONTH
        d.limits = int(run.limits)
        self._desc = d
        ptr = ctypes.c_void_p()
        _lib.check(_lib.lib().tm5_tree_create(D.context().ptr, ctypes.byref(d), ctypes.byref(ptr)), "tm5_tree_create")
        self.ptr = ptr

    def option(self, which, value):
        """tm5_tree_option: _lib.TREE_GRAPHS (capture and replay steps, 1) or _lib.TREE_BATCH."""
        _lib.check(_lib.lib().tm5_tree_option(self.ptr, int(which), int(value)), "tm5_tree_option")

    def interval(self, t0, nread=3600, ndyn_max=3600):
        """tm5_tree_interval: Check_CFL (its verdict read once per attempt), then the steps."""
        nd, nn = ctypes.c_int(), ctypes.c_int()
        rc = _lib.lib().tm5_tree_interval(self.ptr, float(t0), int(nread), int(ndyn_max), ctypes.byref(nd),
                                          ctypes.byref(nn))  # fmt: skip
        if rc == _lib.EFLAGS:
            D.context().raise_errors()
        _lib.check(rc, "tm5_tree_interval")
        return nd.value, nn.value

    def run(self, t0, count, nread=3600, ndyn_max=3600):
        """tm5_tree_run: count intervals queued as checked steps; [(ndyn, n)] per interval."""
        nd, nn = (ctypes.c_int * max(count, 1))(), (ctypes.c_int * max(count, 1))()
        rc = _lib.lib().tm5_tree_run(self.ptr, float(t0), int(nread), int(ndyn_max), int(count), nd, nn)
        if rc == _lib.EFLAGS:
            D.context().raise_errors()
        _lib.check(rc, "tm5_tree_run")
        return [(nd[k], nn[k]) for k in range(count)]

    def tally(self):
        """{'kept', 'redone', 'replayed', 'captured'}: queued intervals kept and redone, steps
        replayed from graphs, graphs captured."""
        v = [ctypes.c_longlong() for _ in range(4)]
        _lib.check(_lib.lib().tm5_tree_tally(self.ptr, *(ctypes.byref(x) for x in v)), "tm5_tree_tally")
        return dict(zip(("kept", "redone", "replayed", "captured"), (x.value for x in v)))

    def close(self):
        if self.ptr:
            _lib.lib().tm5_tree_destroy(self.ptr)
            self.ptr = None


class ZoomRun:
    """The device state, inputs and clock of a two-region tree built from (parent Meteo,
    NestedMeteo)."""

    def __init__(self, pmet, cmet, limits=True, redgrid=None, ensemble=None, driver="native"):
        """ensemble: a pydyn_amd.meteo.Ensemble (its members' tracers side by side, each member's
        emission scaled), or None (the meteo's tracers, the emission into tracer 1).  driver:
        'native' (the library's tm5_tree) or 'python' (pydyn_amd.zoom.driver over the bodies)."""
        if driver not in ("native", "python"):
            raise ValueError(f"ZoomRun: driver {driver!r}: 'native' or 'python'")
        g = pmet.g
        self.mets = {1: pmet, 2: cmet}
        self.zoom = tree_of(g, cmet.c, redgrid)
        self.limits = limits
        self.ensemble = ensemble
        self.driver = driver
        self._native = None
        nt = pmet.nt if ensemble is None else ensemble.nt
        self.st = ZoomState(self.zoom, nt)
        for r, met in self.mets.items():
            init = met.state() if ensemble is None else ensemble.state(met)
            self.st.upload_interior(r, init, met.dxyp if r > 1 else g.dxyp)
        L.update_parent(self.zoom, self.st, 2)  # the footprint holds the child's mass from the start
        self.phys = Physics(self.zoom, self.st, self.mets, g.lmax, None if ensemble is None else ensemble.escale())
        self.flux0 = {r: (D.upload(met.pu0), D.upload(met.pv0)) for r, met in self.mets.items()}
        self.set_fluxes(0.0)
        Z.setup_massflow(self.zoom, self.st, 3600)
        L.m2phlb(self.zoom, self.st, 2)
        L.m2phlb1(self.zoom, self.st, 1)
        self.t = 0.0
        self.log = []

    def set_fluxes(self, t_sec):
        """Every region's pu, pv at time t_sec: the steady flow times the amplitude."""
        a = self.mets[1].amplitude(t_sec)
        for r, (pu0, pv0) in self.flux0.items():
            f = self.st[r]
            L.scaled_copy(pu0, f["pu"].t, a)
            L.scaled_copy(pv0, f["pv"].t, a)

    def native(self):
        """The tree's NativeTree, made on first use."""
        if self._native is None:
            self._native = NativeTree(self)
        return self._native

    def interval(self, nread=3600, ndyn_max=3600):
        """Check_CFL for the tree, then its n steps of ndyn.  Returns (ndyn, n)."""
        t0, zoom, st = self.t, self.zoom, self.st
        if self.driver == "native":
            ndyn, n = self.native().interval(t0, nread, ndyn_max)
            D.context().raise_errors()
            self.t = t0 + n * ndyn
            self.log.append((t0, ndyn, n))
            return ndyn, n

        def interpolate(zoom_, st_, i, ndyn):
            self.set_fluxes(t0 + (i - 0.5) * ndyn)

        ndyn, n, _ = Z.check_cfl_zoom(zoom, st, ndyn_max, nread // ndyn_max, nread, interpolate)
        for s in range(n):
            self.set_fluxes(t0 + s * ndyn + ndyn / 2.0)
            Z.setup_massflow(zoom, st, ndyn)
            Z.step(zoom, st, self.phys, ndyn, self.limits)
        D.context().raise_errors()
        self.t = t0 + n * ndyn
        self.log.append((t0, ndyn, n))
        return ndyn, n

    def run(self, count, nread=3600, ndyn_max=3600):
        """count intervals queued by the native driver (tm5_tree_run: checked steps replayed from
        HIP graphs, the device read once per batch); [(ndyn, n)] per interval, each as
        interval() would give it."""
        out = self.native().run(self.t, count, nread, ndyn_max)
        for ndyn, n in out:
            self.log.append((self.t, ndyn, n))
            self.t = self.t + n * ndyn
        return out

    def host(self):
        return self.st.host()

    def close(self):
        if self._native is not None:
            self._native.close()
            self._native = None
