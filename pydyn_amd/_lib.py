"""ctypes bindings of libtm5hip.so (include/tm5hip.h).

The library is built in-tree by pydyn_amd.build (or __graft_entry__.build()).  There is no
fallback: without the library, or without a GPU, every operator raises.
"""

import ctypes
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libtm5hip.so"

FAIL_X, FAIL_Y, FAIL_Z, BAD_Y, BAD_Z, REHEARSAL, FAIL_XC, FAIL_YC, SPEC = 1, 2, 4, 8, 16, 32, 64, 128, 256
BAD_LEVELS = 512
SINGULAR = 1024
ERRORS = FAIL_X | FAIL_Y | FAIL_Z | BAD_Y | BAD_Z | FAIL_XC | FAIL_YC | BAD_LEVELS | SINGULAR
EINVAL, EFLAGS, ESCRATCH, ENOMEM = -1, -2, -3, -4
NOALLOC_DOUBLES = 4 << 20  # TM5_NOALLOC_DOUBLES: tm5_convec_noalloc's static buffer


class View(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("stride", ctypes.c_int64 * 4)]


class Tracers(ctypes.Structure):
    _fields_ = [
        ("rm", ctypes.c_void_p),
        ("rxm", ctypes.c_void_p),
        ("rym", ctypes.c_void_p),
        ("rzm", ctypes.c_void_p),
        ("stride", ctypes.c_int64 * 4),
        ("nt", ctypes.c_int),
    ]


class IView(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("stride", ctypes.c_int64 * 2)]


class HView(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("stride", ctypes.c_int64 * 4), ("lower", ctypes.c_int32 * 4)]


class ZRegion(ctypes.Structure):
    _fields_ = [
        *((name, ctypes.c_int) for name in ("im", "jm", "lm", "nt", "xref", "yref", "zref", "tref", "tref_abs",
                                            "ibeg", "iend", "jbeg", "jend", "isr", "ier", "jsr", "jer",
                                            "xcyc", "touch_sp", "touch_np", "whole")),
        *((name, HView) for name in ("m", "rm", "rxm", "rym", "rzm", "am", "bm", "cm", "pu", "pv", "sp", "phlb")),
        ("dxyp", ctypes.c_void_p),
        ("owned", ctypes.c_void_p),
        ("edge", ctypes.c_void_p),
        ("nred", ctypes.c_int),
        ("jred", ctypes.c_void_p),
        ("imred", ctypes.c_void_p),
        ("scratch", ctypes.c_void_p),
    ]  # fmt: skip


class ZPhysics(ctypes.Structure):
    _fields_ = [
        *((name, View) for name in ("entu", "detu", "entd", "detd", "dkg")),
        ("cloud_top", IView),
        ("cloud_lfs", IView),
        ("lmax", ctypes.c_int),
        ("dt_conv", ctypes.c_double),
        ("emis", View),
        ("factor", ctypes.c_double),
        ("ratio", ctypes.c_double),
        ("efrac", ctypes.c_double),
        ("level", ctypes.c_int),
        ("tracer", ctypes.c_int),
        ("escale", ctypes.c_void_p),
        ("loss", View),
        ("loss_stride", ctypes.c_int64),
        ("dt_chem", ctypes.c_double),
        ("at", ctypes.c_void_p),
        ("bt", ctypes.c_void_p),
    ]


class RegionDesc(ctypes.Structure):
    _fields_ = [
        ("im", ctypes.c_int),
        ("jm", ctypes.c_int),
        ("lm", ctypes.c_int),
        ("lmax", ctypes.c_int),
        ("limits", ctypes.c_int),
        *((name, View) for name in ("m", "am", "bm", "cm", "pu0", "pv0", "sp", "phlb")),
        ("q", Tracers),
        *((name, View) for name in ("entu", "detu", "entd", "detd", "dkg", "emis", "loss")),
        ("cloud_top", IView),
        ("cloud_lfs", IView),
        ("at", ctypes.c_void_p),
        ("bt", ctypes.c_void_p),
        ("dxyp", ctypes.c_void_p),
        ("mod_period", ctypes.c_double),
        ("mod_amplitude", ctypes.c_double),
        ("sec_month", ctypes.c_double),
        ("mol_ratio", ctypes.c_double),
    ]


class TreeDesc(ctypes.Structure):
    _fields_ = [
        ("nregions", ctypes.c_int),
        ("regions", ctypes.POINTER(ZRegion)),
        ("parent", ctypes.POINTER(ctypes.c_int)),
        ("physics", ctypes.POINTER(ZPhysics)),
        ("pu0", ctypes.POINTER(View)),
        ("pv0", ctypes.POINTER(View)),
        ("at", ctypes.c_void_p),
        ("bt", ctypes.c_void_p),
        ("mod_period", ctypes.c_double),
        ("mod_amplitude", ctypes.c_double),
        ("sec_month", ctypes.c_double),
        ("limits", ctypes.c_int),
    ]


TREE_GRAPHS, TREE_BATCH, TREE_MODE = 0, 1, 2  # tm5_tree_option
TREE_AUTO, TREE_CHECKED, TREE_REHEARSED = 0, 1, 2  # its TREE_MODE values

_P = ctypes.POINTER
_I, _D, _VP = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
_SIGNATURES = {
    "tm5_ctx_create": [ctypes.c_uint64, _P(_VP)],
    "tm5_ctx_destroy": [_VP],
    "tm5_ctx_flags": [_VP, _I, _P(_I)],
    "tm5_ctx_nloop": [_VP, _P(_I), _P(_I), _P(_I), _I, _P(_I)],
    "tm5_ctx_clear": [_VP, _I],
    "tm5_ctx_set_tile": [_VP, _I, _I],
    "tm5_libm": [_VP, _I, ctypes.c_int64, _VP, _VP, _VP],
    "tm5_calc_kzz": [_VP, _I, _I, _I, _I, _P(View), _P(View), _P(View), _P(View), _P(View), _P(View), _P(View), _VP,
                     _VP, _VP, _VP, _D, _P(View), _P(View), _P(View), _P(View), _P(View)],
    "tm5_loss_rates": [_VP, _I, _I, _I, _I, _I, _P(View), _P(View), _VP, _VP, _VP, _VP, _VP, _P(View), _P(View),
                       _P(View), _P(View), _D, _D, _D, _D, _D, _P(View), _P(View)],
    "tm5_ctx_get_tile": [_VP, _I, _P(_I)],
    "tm5_profile_begin": [_VP],
    "tm5_profile_end": [_VP],
    "tm5_profile_entry": [_VP, _I, _P(ctypes.c_char_p), _P(ctypes.c_longlong), _P(_D)],
    "tm5_advect_x": [_VP, _I, _I, _I, _P(View), _P(Tracers), _P(View), _I, _I],
    "tm5_advect_y": [_VP, _I, _I, _I, _P(View), _P(Tracers), _P(View), _I, _I],
    "tm5_advect_z": [_VP, _I, _I, _I, _P(View), _P(Tracers), _P(View), _I, _I],
    "tm5_mass_x": [_VP, _I, _I, _I, _P(View), _P(View), _I],
    "tm5_mass_y": [_VP, _I, _I, _I, _P(View), _P(View), _I],
    "tm5_mass_z": [_VP, _I, _I, _I, _P(View), _P(View), _I],
    "tm5_dynam0": [_VP, _I, _I, _I, _P(View), _P(View), _D, _VP, _D, _P(View), _P(View), _P(View)],
    "tm5_m2phlb1": [_VP, _I, _I, _I, _P(View), _VP, _VP, _VP, _P(View), _P(View)],
    "tm5_convec": [_VP, _I, _I, _I, _P(View), _P(Tracers), *([_P(View)] * 4), _P(IView), _P(IView), _P(View), _D, _I],
    "tm5_dkg": [_VP, _I, _I, _P(View), _P(View), _P(View), _I, _P(View)],
    "tm5_add_2d": [_VP, _I, _I, _P(View), _D, _D, _D, _P(Tracers), _I, _I],
    "tm5_decay": [_VP, _I, _I, _I, _P(View), _D, _P(Tracers), _I, _I],
    "tm5_scale": [_VP, _I, _I, _I, _P(View), _D],
    "tm5_region_create": [_VP, _P(RegionDesc), _P(_VP)],
    "tm5_region_destroy": [_VP],
    "tm5_region_interval": [_VP, _D, _I, _I, _P(_I), _P(_I)],
    "tm5_advect_x_rows": [_VP, _I, _I, _I, _P(View), _P(Tracers), _P(View), _VP, _I, _I, _VP, _VP],
    "tm5_mass_x_rows": [_VP, _I, _I, _I, _P(View), _P(View), _VP, _I],
    "tm5_advect_z_substeps": [_VP, _I, _I, _I, _P(View), _P(Tracers), _P(View), _I, _I, _I],
    "tm5_cluster_sum": [_VP, _I, _I, _I, _I, _P(View), _P(View)],
    "tm5_convec_noalloc": [_VP, _I, _I, _I, _P(View), _P(Tracers), *([_P(View)] * 4), _P(IView), _P(IView), _P(View),
                           _D, _I],
    "tm5_convec_masked": [_VP, _I, _I, _I, _P(View), _P(Tracers), _P(View), _P(View), _P(View), _P(View), _P(IView),
                          _P(IView), _P(View), _D, _I, _VP],
    "tm5_convec_lapack": [_VP, _I, _I, _I, _P(View), _P(Tracers), _P(View), _P(View), _P(View), _P(View), _P(IView),
                          _P(IView), _P(View), _D, _I, _VP],
    "tm5_inverse_rows": [_VP, ctypes.c_longlong, _I, _VP, _VP, _VP],
    "tm5_region_ops": [_VP, _D, _I, ctypes.c_char_p],
    "tm5_region_run": [_VP, _D, _I, _I, _I, _P(_I), _P(_I)],
    "tm5_advect_adjoint": [_VP, _I, _I, _I, _I, _P(View), _P(Tracers), _P(View), _P(Tracers), _I, _I, _VP],
    "tm5_zoom_put_xedges": [_VP, _P(ZRegion), _P(ZRegion), _I],
    "tm5_zoom_put_yedges": [_VP, _P(ZRegion), _P(ZRegion), _I],
    "tm5_zoom_compress_xedges": [_VP, _P(ZRegion), _I, _I, _I, _I, _I],
    "tm5_zoom_uncompress_xedges": [_VP, _P(ZRegion), _I, _I, _I, _I, _I],
    "tm5_zoom_compress_yedges": [_VP, _P(ZRegion), _I, _I, _I, _I, _I],
    "tm5_zoom_uncompress_yedges": [_VP, _P(ZRegion), _I, _I, _I, _I, _I],
    "tm5_zoom_mix_edges": [_VP, _P(ZRegion), _I],
    "tm5_zoom_update_parent": [_VP, _P(ZRegion), _P(ZRegion), _I],
    "tm5_zoom_coarsen_block": [_VP, _P(ZRegion), _P(ZRegion)],
    "tm5_zoom_coarsen_ambm": [_VP, _P(ZRegion), _P(ZRegion)],
    "tm5_zoom_m2phlb": [_VP, _P(ZRegion), _P(ZRegion), _VP, _VP],
    "tm5_zoom_dynam0": [_VP, _P(ZRegion), _VP, _I, _I],
    "tm5_zoom_outer_faces": [_VP, _P(ZRegion), _I, _I, _VP, _I],
    "tm5_zoom_advect_x": [_VP, _P(ZRegion), _I, _I, _I, _I, _I],
    "tm5_zoom_advect_y": [_VP, _P(ZRegion), _I, _I, _I, _I, _I],
    "tm5_zoom_advect_z": [_VP, _P(ZRegion), _I, _I, _I, _I, _I, _I, _I],
    "tm5_zoom_physics": [_VP, _P(ZRegion), _P(ZPhysics), ctypes.c_char_p],
    "tm5_zoom_copy": [_VP, _VP, _VP, ctypes.c_int64],
    "tm5_zoom_inject": [_VP, _VP, _P(View), _I, _I],
    "tm5_zoom_decay": [_VP, _VP, _P(View), _I, _I],
    "tm5_column_adjoint": [_VP, _I, _I, _I, _P(View), _P(Tracers), _P(ZPhysics), ctypes.c_char_p, _P(View)],
    "tm5_zoom_scale": [_VP, _VP, ctypes.c_int64, _D],
    "tm5_tree_create": [_VP, _P(TreeDesc), _P(_VP)],
    "tm5_tree_destroy": [_VP],
    "tm5_tree_interval": [_VP, _D, _I, _I, _P(_I), _P(_I)],
    "tm5_tree_run": [_VP, _D, _I, _I, _I, _P(_I), _P(_I)],
    "tm5_tree_option": [_VP, _I, _I],
    "tm5_tree_tally": [_VP, *([_P(ctypes.c_longlong)] * 4)],
}

_STATE = {}


class LibraryMissing(RuntimeError):
    """libtm5hip.so is not built (run pydyn_amd.build or __graft_entry__.build())."""


def lib():
    """The loaded library, its prototypes set; raises LibraryMissing when it is not built."""
    if "lib" not in _STATE:
        if not LIB_PATH.exists():
            raise LibraryMissing(f"{LIB_PATH} is not built: run `python -m pydyn_amd.build`")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, args in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype, fn.argtypes = ctypes.c_int, args
        _STATE["lib"] = handle
    return _STATE["lib"]


def check(rc, what):
    if rc == ESCRATCH:
        raise ValueError(f"libtm5hip: {what}: one row of columns needs more than the static buffer's "
                         f"{NOALLOC_DOUBLES} doubles (im lmax^2 too large)")
    if rc == ENOMEM:
        raise MemoryError(f"libtm5hip: {what}: host memory exhausted")
    if rc != 0:
        raise RuntimeError(f"libtm5hip: {what} failed with code {rc}")
