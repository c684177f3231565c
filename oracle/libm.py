"""The oracle's exp, log and pow: the fdlibm methods in C (oracle/libm.c), built into
oracle/_build/ by oracle/Makefile, applied element by element through a C loop.  Test
infrastructure: the checker of the device's libm, never a product path."""

import ctypes
import pathlib
import subprocess

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "_build" / "liboracle_libm.so"
_LIB = []


def build():
    """Compile oracle/libm.c (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", str(LIB_PATH.parent.parent)], check=True)


def _lib():
    if not _LIB:
        if not LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(LIB_PATH))
        for name in ("oracle_exp_map", "oracle_log_map"):
            getattr(lib, name).restype = None
            getattr(lib, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        lib.oracle_pow_map.restype = None
        lib.oracle_pow_map.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        _LIB.append(lib)
    return _LIB[0]


def exp(x):
    """fdlibm exp of every element of x (float64)."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    y = np.empty_like(x)
    _lib().oracle_exp_map(x.ctypes.data, y.ctypes.data, x.size)
    return y


def log(x):
    """fdlibm log of every element of x (float64)."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    y = np.empty_like(x)
    _lib().oracle_log_map(x.ctypes.data, y.ctypes.data, x.size)
    return y


def pow(x, y):  # noqa: A001 -- the libm name
    """fdlibm pow of every pair of elements (float64, broadcast)."""
    x, y = np.broadcast_arrays(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64))
    x, y = np.ascontiguousarray(x), np.ascontiguousarray(y)
    out = np.empty_like(x)
    _lib().oracle_pow_map(x.ctypes.data, y.ctypes.data, out.ctypes.data, x.size)
    return out
