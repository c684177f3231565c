"""The bodies the zoom driver runs on a device state: pydyn.gpu.zoom's names and arguments
(pydyn/gpu/zoom.py:38-74, the list of pydyn/zoom/host.py:42-48), each a call into libtm5hip,
and the state's zeros / upload / download and the queued submission Pipeline (pydyn/gpu/zoom.py:
77-101, 917-1041).

Every body is asynchronous on the context's stream.  The rehearsal's mass-only bodies do not
return a verdict: a failure sets TM5_REHEARSAL in the context's flag word, after which every
later rehearsal kernel does nothing, so the driver reads the verdict once per Check_CFL
iteration.  The main bodies' errors (CFLError / ValueError) are read by Context.raise_errors.
"""

import ctypes
from contextlib import contextmanager

from .. import _lib
from .. import device as D
from ..dump import FArray
from .state import HostState, ZoomState

__all__ = [
    "Pipeline", "_coarsen_block", "_dynamu_child", "_dynamu_region1", "_dynamum_child", "_dynamum_region1",
    "_dynamv_child", "_dynamv_region1", "_dynamvm_child", "_dynamvm_region1", "_dynamw_child",
    "_dynamw_region1", "_dynamwm", "_uncompress_xedges", "_uncompress_yedges", "assign", "chemistry_region",
    "coarsen_ambm", "column_physics", "convec_region", "download", "dynam0", "m2phlb", "m2phlb1", "mix_edges",
    "outer_x_faces_zeroed", "outer_y_faces_zeroed", "put_xedges", "put_yedges", "restore", "save", "scale",
    "sources_region", "update_parent", "upload", "zeros",
]  # fmt: skip


# ------------------------------------------------------------------ the state on the device
def zeros(zoom, nt):
    """Every region's arrays on the device, zero (ZoomState.zeros there; pydyn/gpu/zoom.py:77-88)."""
    return ZoomState(zoom, nt)


def upload(st, tag="seed", zoom=None):
    """A host state -- fields[r][name] with .data in the declared shape and .lower, as download
    returns it or the reference's host ZoomState holds it -- on the device (pydyn/gpu/zoom.py:
    91-98).  zoom: the tree the state belongs to (default st.zoom).  tag is the reference's
    transfer label and changes nothing."""
    zoom = zoom if zoom is not None else getattr(st, "zoom", None)
    if zoom is None:
        raise ValueError("upload: the state's Zoom is needed (pass zoom=)")
    dev = ZoomState(zoom, st.fields[1]["rm"].data.shape[3])
    for r, fields in st.fields.items():
        for k, v in fields.items():
            f = dev[r][k]
            if tuple(v.lower) != f.lower or tuple(v.data.shape) != tuple(f.t.shape):
                raise ValueError(f"upload: region {r} {k}: bounds {tuple(v.lower)} shape {v.data.shape}, "
                                 f"declared {f.lower} {tuple(f.t.shape)}")  # fmt: skip
            f.t.copy_(D.upload(v.data))
    return dev


def download(st, tag="verify"):
    """The device state's arrays on the host as a HostState of FArrays (pydyn/gpu/zoom.py:
    101-104); tag as upload's."""
    return HostState({r: {k: FArray(v.host(), v.lower) for k, v in f.items()} for r, f in st.fields.items()},
                     st.zoom)  # fmt: skip


# ------------------------------------------------------------------ queued submission
def _interval_of(fn):
    """The ZoomRun whose own interval fn is (submit(run.interval)), else None."""
    from .run import ZoomRun

    owner = getattr(fn, "__self__", None)
    ok = isinstance(owner, ZoomRun) and owner.driver == "native" and getattr(fn, "__name__", "") == "interval"
    return owner if ok else None


class Pipeline:
    """Calls fn() -- any sequence of the driver's bodies on st -- in submission order, up to
    `depth` in flight (pydyn/gpu/zoom.py:917-1041); drain() returns the calls' results.

    A ZoomRun's own interval (submit(run.interval, repeat=...)) goes to the library's queue
    (tm5_tree_run): each interval planned on the previous one's step, its steps checked on the
    device or its planned trial rehearsed and verified there, the host reading the device once per
    `depth` intervals, an interval whose plan fails redone the reference's way -- the results are
    always the careful run's bits.  With `repeat` given, every step of a given length is captured
    into a HIP graph once and replayed (its time comes from a device clock, so every interval
    replays the same graph).  Any other fn runs when submitted: the bodies are asynchronous on the
    context's stream and read the device only where the reference's careful run decides (Check_CFL's
    verdict).  `tally`: the intervals kept and redone, and the steps replayed from graphs."""

    def __init__(self, zoom, st, depth=2):
        if depth < 1:
            raise ValueError("Pipeline: depth >= 1")
        self.zoom, self.st, self.depth = zoom, st, depth
        self.done = []
        self.tally = {"kept": 0, "redone": 0, "replayed": 0}
        self._pending = None  # [run, intervals, repeat]

    def submit(self, fn, repeat=None):
        run = _interval_of(fn)
        if run is None:
            self._flush()
            self.done.append(fn())
            self.tally["kept"] += 1
            return
        if self._pending is not None and (self._pending[0] is not run or self._pending[2] != repeat):
            self._flush()
        if self._pending is None:
            self._pending = [run, 0, repeat]
        self._pending[1] += 1
        if self._pending[1] >= self.depth:
            self._flush()

    def _flush(self):
        if self._pending is None:
            return
        run, count, repeat = self._pending
        self._pending = None
        tree = run.native()
        tree.option(_lib.TREE_GRAPHS, 0 if repeat is None else 1)
        before = tree.tally()
        self.done.extend(run.run(count))
        after = tree.tally()
        for k in ("kept", "redone", "replayed"):
            self.tally[k] += after[k] - before[k]

    def drain(self):
        """Settle every call in flight (raising the reference's exception for an error the device
        recorded); the results so far."""
        self._flush()
        D.context().raise_errors()
        out, self.done = self.done, []
        return out


def _call(name, *args):
    _lib.check(getattr(_lib.lib(), name)(D.context().ptr, *args), name)


# ------------------------------------------------------------------ exchange
def put_xedges(zoom, st, region, tracers=True):
    for c in zoom.children[region]:
        _call("tm5_zoom_put_xedges", st.ptr(region), st.ptr(c), int(tracers))


def put_yedges(zoom, st, region, tracers=True):
    for c in zoom.children[region]:
        _call("tm5_zoom_put_yedges", st.ptr(region), st.ptr(c), int(tracers))


def _uncompress_xedges(zoom, st, region, js, je, ls, le, tracers=True):
    _call("tm5_zoom_uncompress_xedges", st.ptr(region), js, je, ls, le, int(tracers))


def _uncompress_yedges(zoom, st, region, is_, ie, ls, le, tracers=True):
    _call("tm5_zoom_uncompress_yedges", st.ptr(region), is_, ie, ls, le, int(tracers))


def mix_edges(zoom, st, region, tracers=True):
    _call("tm5_zoom_mix_edges", st.ptr(region), int(tracers))


def update_parent(zoom, st, region, tracers=True):
    if region != 1:
        _call("tm5_zoom_update_parent", st.ptr(zoom.g(region).parent), st.ptr(region), int(tracers))


# ------------------------------------------------------------------ mass flow
def m2phlb1(zoom, st, region=1):
    _call("tm5_zoom_m2phlb", st.ptr(region), None, st.at.data_ptr(), st.bt.data_ptr())


def m2phlb(zoom, st, region):
    _call("tm5_zoom_m2phlb", st.ptr(region), st.ptr(zoom.g(region).parent), st.at.data_ptr(), st.bt.data_ptr())


def _coarsen_block(zoom, st, region):
    _call("tm5_zoom_coarsen_block", st.ptr(zoom.g(region).parent), st.ptr(region))


def coarsen_ambm(zoom, st, region):
    for c in zoom.children[region]:
        _call("tm5_zoom_coarsen_ambm", st.ptr(region), st.ptr(c))


def dynam0(zoom, st, region, ndyn, rehearsal=False):
    _call("tm5_zoom_dynam0", st.ptr(region), st.bt.data_ptr(), int(ndyn), int(rehearsal))


# ------------------------------------------------------------------ outer interface faces
@contextmanager
def outer_x_faces_zeroed(zoom, st, region, second, rehearsal=False):
    """A child's outer x interface faces zeroed around its x advection in the second half of the
    parent step (pydyn/zoom/advect.py:121-136)."""
    g = zoom.g(region)
    zero = second and g.im // g.xref < zoom.g(1).im
    buf = st.slabs[region][0].data_ptr()
    if zero:
        _call("tm5_zoom_outer_faces", st.ptr(region), 0, 0, buf, int(rehearsal))
    yield
    if zero:
        _call("tm5_zoom_outer_faces", st.ptr(region), 0, 1, buf, int(rehearsal))


@contextmanager
def outer_y_faces_zeroed(zoom, st, region, second, rehearsal=False):
    """The same for the y interface faces of a region touching no pole (advect.py:139-157)."""
    g = zoom.g(region)
    zero = second and (g.touch_sp != 1 or g.touch_np != 1)
    buf = st.slabs[region][1].data_ptr()
    if zero:
        _call("tm5_zoom_outer_faces", st.ptr(region), 1, 0, buf, int(rehearsal))
    yield
    if zero:
        _call("tm5_zoom_outer_faces", st.ptr(region), 1, 1, buf, int(rehearsal))


# ------------------------------------------------------------------ advection
def _dynamu_region1(zoom, st, js, je, limits, stats=None, max_nloop=50):
    _call("tm5_zoom_advect_x", st.ptr(1), js, je, int(limits), max_nloop, 0)


def _dynamu_child(zoom, st, region, js, je, ls, le, limits, stats=None, max_nloop=50):
    if (ls, le) != (1, zoom.lm):
        raise NotImplementedError("_dynamu_child: zoom2D levels only")
    _call("tm5_zoom_advect_x", st.ptr(region), js, je, int(limits), max_nloop, 0)


def _dynamv_region1(zoom, st, limits, stats=None, max_nloop=10):
    _call("tm5_zoom_advect_y", st.ptr(1), 1, zoom.g(1).im, int(limits), max_nloop, 0)


def _dynamv_child(zoom, st, region, is_, ie, ls, le, limits, stats=None, max_nloop=10):
    if (ls, le) != (1, zoom.lm):
        raise NotImplementedError("_dynamv_child: zoom2D levels only")
    _call("tm5_zoom_advect_y", st.ptr(region), is_, ie, int(limits), max_nloop, 0)


def _dynamw_region1(zoom, st, limits, stats=None, iter_limit=15):
    g = zoom.g(1)
    _call("tm5_zoom_advect_z", st.ptr(1), 1, g.im, 1, g.jm, int(limits), iter_limit, 0)


def _dynamw_child(zoom, st, region, is_, ie, js, je, limits, stats=None, iter_limit=15):
    _call("tm5_zoom_advect_z", st.ptr(region), is_, ie, js, je, int(limits), iter_limit, 0)


# the rehearsal's mass-only forms (pydyn/zoom/check_cfl.py:101-226)
def _dynamum_region1(zoom, st, js, je, max_nloop):
    _call("tm5_zoom_advect_x", st.ptr(1), js, je, 0, max_nloop, 1)


def _dynamum_child(zoom, st, region, js, je, ls, le, max_nloop):
    _call("tm5_zoom_advect_x", st.ptr(region), js, je, 0, max_nloop, 1)


def _dynamvm_region1(zoom, st, max_nloop):
    _call("tm5_zoom_advect_y", st.ptr(1), 1, zoom.g(1).im, 0, max_nloop, 1)


def _dynamvm_child(zoom, st, region, is_, ie, ls, le, max_nloop):
    _call("tm5_zoom_advect_y", st.ptr(region), is_, ie, 0, max_nloop, 1)


def _dynamwm(zoom, st, region, is_, ie, js, je, iter_limit):
    _call("tm5_zoom_advect_z", st.ptr(region), is_, ie, js, je, 0, iter_limit, 1)


# ------------------------------------------------------------------ column physics
def column_physics(zoom, st, region, phys, ops):
    """convec with the sources and chemistry around it (pydyn/zoom/physics.py:34-78) in one
    pass over the region's owned columns; `ops` one of v, csv, vsc, vsccsv, vscMcsv."""
    _call("tm5_zoom_physics", st.ptr(region), ctypes.byref(phys), ops.encode())


def convec_region(zoom, st, region, conv, ndyn, inverse="crout"):
    """convec on the columns the region owns (pydyn/zoom/physics.py:33-52), dt = ndyn / (2 tref):
    conv holds device arrays entu, detu, entd, detd, dkg (im, jm, >= lmax) and cloud_top,
    cloud_lfs (im, jm) int32, plus 'lmax'.  inverse: 'crout' only."""
    if inverse != "crout":
        raise ValueError(f"convec_region: inverse {inverse!r}: only 'crout' (fastminv) is provided")
    p = _lib.ZPhysics()
    for k in ("entu", "detu", "entd", "detd", "dkg"):
        setattr(p, k, D.view(conv[k]))
    p.cloud_top, p.cloud_lfs = D.iview(conv["cloud_top"]), D.iview(conv["cloud_lfs"])
    p.lmax = int(conv["lmax"])
    p.dt_conv = float(ndyn) / (2.0 * zoom.g(region).tref)
    _call("tm5_zoom_physics", st.ptr(region), ctypes.byref(p), b"v")


def sources_region(zoom, st, region, injections):
    """do_add_2d on the cells the region owns (pydyn/zoom/physics.py:55-64): each (x (im, jm)
    device array in kg, level, tracer), 1-based, added to rm and subtracted from rzm."""
    for x, level, tracer in injections:
        _call("tm5_zoom_inject", st.ptr(region), ctypes.byref(D.view(x)), int(level), int(tracer))


def chemistry_region(zoom, st, region, decay, tracers):
    """rm and the slopes of the 1-based `tracers` times `decay` (the scope's factors, a device
    array (ier-isr+1, jer-jsr+1, lm)) on the owned cells (pydyn/zoom/physics.py:67-78)."""
    v = D.view(decay)
    for t in tracers:
        _call("tm5_zoom_decay", st.ptr(region), ctypes.byref(v), int(t), 1)


def assign(dst, src, region, names):
    """dst's named arrays of the region set to src's (pydyn/zoom/host.py:51-54)."""
    for k in names:
        a, b = src[region][k], dst[region][k]
        _call("tm5_zoom_copy", a.t.data_ptr(), b.t.data_ptr(), b.numel())


# ------------------------------------------------------------------ save / restore / scale
def save(st, key, names):
    """Copies of the named arrays of every region, for restore(st, key)."""
    bufs = st.saved.setdefault(key, {})
    for r, f in st.fields.items():
        for k in names:
            if (r, k) not in bufs:
                bufs[(r, k)] = D.empty((f[k].numel(),))
            _call("tm5_zoom_copy", f[k].t.data_ptr(), bufs[(r, k)].data_ptr(), f[k].numel())


def restore(st, key):
    for (r, k), buf in st.saved[key].items():
        f = st.fields[r][k]
        _call("tm5_zoom_copy", buf.data_ptr(), f.t.data_ptr(), f.numel())


def scale(st, names, factor):
    """a = a * factor for the named arrays of every region."""
    for f in st.fields.values():
        for k in names:
            _call("tm5_zoom_scale", f[k].t.data_ptr(), f[k].numel(), float(factor))


def scaled_copy(src, dst, factor):
    """dst = src * factor (whole arrays): a meteo field at amplitude `factor`."""
    _call("tm5_zoom_copy", src.data_ptr(), dst.data_ptr(), dst.numel())
    _call("tm5_zoom_scale", dst.data_ptr(), dst.numel(), float(factor))
