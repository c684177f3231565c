"""The oracle's LAPACK-kind inverse (oracle/lapack.c: reference LAPACK's DGETRF2 + DGETRI with
lwork = n, as pydyn/lapack.py:43-66 calls them per column), and -- where oracle/_ref is built
(make -C oracle ref, from the reference's own sources) -- the reference's twin of the same
functions, the checker that pins the restatement.  Test infrastructure only."""

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle_lapack.so"
REF_PATH = HERE / "_ref" / "libpd_host.so"
REF_SOURCES = pathlib.Path("/root/reference/pydyn/gpu/csrc")
_LIB = {}


def build():
    """Compile oracle/lapack.c; and the reference's host build when its sources are here."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    if REF_SOURCES.is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref"], check=True)


def _lib():
    if "oracle" not in _LIB:
        if not LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(LIB_PATH))
        lib.oracle_inverse_rows.restype = ctypes.c_int
        lib.oracle_inverse_rows.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        _LIB["oracle"] = lib
    return _LIB["oracle"]


def inverse_rows(a):
    """The inverse of each (n, n) matrix of a (ncol, n, n), row-major in and out.  Raises
    ValueError when a pivot is exactly zero (the reference's LapackInverse does)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    ncol, n, _ = a.shape
    y = np.empty_like(a)
    bad = _lib().oracle_inverse_rows(ncol, n, a.ctypes.data, y.ctypes.data)
    if bad:
        raise ValueError(f"inverse_rows: {bad} singular matrices")
    return y


def reference_available():
    return REF_PATH.exists()


def reference_inverse_rows(a):
    """The same through the reference's dgetrf_ / dgetri_ twin (oracle/_ref), one column-major
    copy per matrix as pydyn/lapack.py:50-66 does."""
    if "ref" not in _LIB:
        lib = ctypes.CDLL(str(REF_PATH))
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.c_void_p
        lib.dgetrf_.argtypes = [ip, ip, dp, ip, dp, ip]
        lib.dgetri_.argtypes = [ip, dp, ip, dp, dp, ip, ip]
        _LIB["ref"] = lib
    lib = _LIB["ref"]
    a = np.asarray(a, dtype=np.float64)
    ncol, n, _ = a.shape
    cn, info = ctypes.c_int(n), ctypes.c_int(0)
    ipiv, work = np.zeros(n, dtype=np.int32), np.zeros(n)
    out = np.empty_like(a)
    for c in range(ncol):
        af = np.asfortranarray(a[c])
        lib.dgetrf_(ctypes.byref(cn), ctypes.byref(cn), af.ctypes.data, ctypes.byref(cn), ipiv.ctypes.data, ctypes.byref(info))
        if info.value:
            raise ValueError(f"dgetrf: info {info.value} for matrix {c}")
        lib.dgetri_(ctypes.byref(cn), af.ctypes.data, ctypes.byref(cn), ipiv.ctypes.data, work.ctypes.data,
                    ctypes.byref(cn), ctypes.byref(info))  # fmt: skip
        if info.value:
            raise ValueError(f"dgetri: info {info.value} for matrix {c}")
        out[c] = af
    return out
