"""Device arrays and the library context.

Arrays live in HBM as torch tensors (torch is the allocator and stream provider here, nothing
else): a logical pydyn-layout shape (im, jm, lm[, nt]) stored with i fastest -- the physical
order is the reversed axes -- so that every sweep of the kernels reads consecutive doubles
across lanes.  `upload` / `download` convert from / to C-ordered NumPy arrays in the pydyn
layout.  Every operator runs on the context's stream (torch's current stream when the context
was made), so torch copies and library kernels are ordered on it.
"""

import ctypes
import json
import pathlib

import numpy as np

from . import _lib

# tm5_ctx_set_tile's knobs (include/tm5hip.h), by name
TILES = {"y_lines": 0, "z_lines": 1, "dynam0": 2, "m2phlb": 3, "y_many": 4, "range_min": 5, "apply": 6}
# the shapes pydyn_amd.tune measured best, applied to every new context (absent: the defaults)
TILES_FILE = pathlib.Path(__file__).resolve().parent / "tiles.json"
TILES_FILE_ENABLED = True


class CFLError(RuntimeError):
    """A sub-cycle count exceeded its cap (the reference's pydyn.kernels.CFLError)."""


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("pydyn_amd needs an AMD Instinct GPU with ROCm: torch.cuda.is_available() is False")
    return torch


class Context:
    """A tm5_ctx bound to one stream: the flag word and the search scratch of the operators."""

    def __init__(self, stream=None):
        torch = _torch()
        lib = _lib.lib()
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        ptr = ctypes.c_void_p()
        _lib.check(lib.tm5_ctx_create(self.stream.cuda_stream, ctypes.byref(ptr)), "tm5_ctx_create")
        self.ptr = ptr
        if TILES_FILE_ENABLED and TILES_FILE.exists():
            for name, value in json.loads(TILES_FILE.read_text())["tiles"].items():
                self.set_tile(name, value)

    def set_tile(self, name, value):
        """A tile shape of the kernels (TILES; speed only, never results)."""
        _lib.check(_lib.lib().tm5_ctx_set_tile(self.ptr, TILES[name], int(value)), f"tm5_ctx_set_tile {name}")

    def tile(self, name):
        out = ctypes.c_int()
        _lib.check(_lib.lib().tm5_ctx_get_tile(self.ptr, TILES[name], ctypes.byref(out)), "tm5_ctx_get_tile")
        return out.value

    def flags(self, clear=True):
        """Wait for the stream and return the flag word (cleared when `clear`)."""
        out = ctypes.c_int()
        _lib.check(_lib.lib().tm5_ctx_flags(self.ptr, int(clear), ctypes.byref(out)), "tm5_ctx_flags")
        return out.value

    def raise_errors(self):
        """The reference's exceptions for the error bits of the flag word (clearing it)."""
        f = self.flags(clear=True)
        if f & _lib.BAD_Y:
            raise ValueError("advect_y: non-zero flux through the polar outer faces")
        if f & _lib.BAD_Z:
            raise ValueError("dynamw: non-zero halo interfaces in cm")
        if f & _lib.FAIL_X:
            raise CFLError("dynamu: nloop too high")
        if f & _lib.FAIL_Y:
            raise CFLError("dynamv: nloop too high")
        if f & _lib.FAIL_Z:
            raise CFLError("dynamw: too many CFL violations!")
        if f & _lib.FAIL_XC:
            raise CFLError("dynuamu: nloop too high")  # the reference's spelling (pydyn/zoom/advect.py:330)
        if f & _lib.FAIL_YC:
            raise CFLError("nloop too high in dynamv")
        if f & _lib.BAD_LEVELS:
            raise ValueError("rebin: levels not strictly descending (the Fortran keeps the stale column)")
        if f & _lib.SINGULAR:
            raise ValueError("pydyn.gpu.convec: a singular convection matrix (dgetrf / dgetri info != 0)")
        return f

    def nloop(self, lm):
        """(x max nloop, y per-layer nloop (lm,), z n_advz) of the last operators."""
        x, y, z = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        layers = (ctypes.c_int * max(lm, 1))()
        _lib.check(_lib.lib().tm5_ctx_nloop(self.ptr, ctypes.byref(x), ctypes.byref(y), layers, lm, ctypes.byref(z)),
                   "tm5_ctx_nloop")  # fmt: skip
        return x.value, np.array(layers[:lm]), z.value

    def profile_begin(self):
        """Time every launch on the context's stream until profile_end (pydyn/gpu/backend.py:633)."""
        _lib.check(_lib.lib().tm5_profile_begin(self.ptr), "tm5_profile_begin")

    def profile_end(self):
        """{kernel: [launches, seconds]} since profile_begin, from events on the stream
        (pydyn/gpu/backend.py:639-649)."""
        _lib.check(_lib.lib().tm5_profile_end(self.ptr), "tm5_profile_end")
        return self.profile_entries()

    def profile_entries(self):
        """The totals of the last profile_end."""
        lib = _lib.lib()
        out, idx = {}, 0
        name, count, secs = ctypes.c_char_p(), ctypes.c_longlong(), ctypes.c_double()
        while lib.tm5_profile_entry(self.ptr, idx, ctypes.byref(name), ctypes.byref(count), ctypes.byref(secs)) == 0:
            out[name.value.decode()] = [count.value, secs.value]
            idx += 1
        return out

    def close(self):
        if self.ptr:
            _lib.lib().tm5_ctx_destroy(self.ptr)
            self.ptr = None


_CONTEXTS = {}


def context():
    """This process's context on the current device."""
    torch = _torch()
    dev = torch.cuda.current_device()
    if dev not in _CONTEXTS:
        _CONTEXTS[dev] = Context()
    return _CONTEXTS[dev]


# ------------------------------------------------------------------ arrays
def empty(shape, dtype=None):
    """An uninitialised device array of logical shape `shape`, i fastest."""
    torch = _torch()
    dtype = torch.float64 if dtype is None else dtype
    t = torch.empty(tuple(reversed(shape)), dtype=dtype, device="cuda")
    return t.permute(*reversed(range(len(shape))))


def zeros(shape, dtype=None):
    t = empty(shape, dtype)
    t.zero_()
    return t


def upload(a, dtype=None):
    """A C-ordered NumPy array (pydyn layout) as a device array, i fastest."""
    torch = _torch()
    a = np.asarray(a)
    host = np.ascontiguousarray(np.transpose(a))
    t = torch.from_numpy(host).to("cuda")
    if dtype is not None:
        t = t.to(dtype)
    return t.permute(*reversed(range(a.ndim)))


def download(t):
    """A device array as a C-ordered NumPy array in its logical shape."""
    host = t.permute(*reversed(range(t.dim()))).contiguous().cpu().numpy()
    return np.ascontiguousarray(np.transpose(host))


def view(t):
    """tm5_view of a float64 device array (up to 4 axes)."""
    if t.dtype.is_floating_point is False or t.element_size() != 8:
        raise TypeError("view: float64 device arrays only")
    strides = list(t.stride()) + [0] * (4 - t.dim())
    return _lib.View(t.data_ptr(), (ctypes.c_int64 * 4)(*strides))


def hview(t, lower):
    """tm5_hview of a float64 device array with Fortran lower bounds `lower` (3 or 4 axes; a
    3-axis array gets a unit t axis at lower bound 1)."""
    if t.dtype.is_floating_point is False or t.element_size() != 8:
        raise TypeError("hview: float64 device arrays only")
    strides = list(t.stride()) + [0] * (4 - t.dim())
    lo = list(lower) + [1] * (4 - len(lower))
    return _lib.HView(t.data_ptr(), (ctypes.c_int64 * 4)(*strides), (ctypes.c_int32 * 4)(*lo))


def iview(t):
    if t.element_size() != 4 or t.dtype.is_floating_point:
        raise TypeError("iview: int32 device arrays only")
    return _lib.IView(t.data_ptr(), (ctypes.c_int64 * 2)(*t.stride()[:2]))


def tracers(rm, rxm, rym, rzm):
    """tm5_tracers of four (im, jm, lm, nt) arrays with one stride set."""
    for q in (rxm, rym, rzm):
        if q.shape != rm.shape or q.stride() != rm.stride():
            raise ValueError("tracers: rm, rxm, rym, rzm must share shape and strides")
    strides = list(rm.stride()) + [0] * (4 - rm.dim())
    nt = rm.shape[3] if rm.dim() == 4 else 1
    return _lib.Tracers(rm.data_ptr(), rxm.data_ptr(), rym.data_ptr(), rzm.data_ptr(),
                        (ctypes.c_int64 * 4)(*strides), nt)  # fmt: skip
