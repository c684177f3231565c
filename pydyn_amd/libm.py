"""The device's libm over device arrays: the reference's `ops.libm` route (pydyn/ops.py:165-232)
for exp, log and pow, by the fdlibm methods every kernel of the library uses (tm5_libm).

    from pydyn_amd import libm
    y = libm.exp(x)          # x a float64 device tensor of any shape
    z = libm.pow(x, y)       # same shapes
"""

from . import _lib
from . import device as D

_FN = {"exp": 0, "log": 1, "pow": 2}


def _map(fn, x, y=None):
    torch = D._torch()
    x = x.to(torch.float64).contiguous()
    if y is not None:
        y = y.to(torch.float64).contiguous()
        if y.shape != x.shape:
            raise ValueError(f"libm.pow: shapes {tuple(x.shape)} and {tuple(y.shape)} differ")
    out = torch.empty_like(x)
    ctx = D.context()
    rc = _lib.lib().tm5_libm(ctx.ptr, _FN[fn], x.numel(), x.data_ptr(), 0 if y is None else y.data_ptr(), out.data_ptr())
    _lib.check(rc, f"tm5_libm {fn}")
    return out


def exp(x):
    return _map("exp", x)


def log(x):
    return _map("log", x)


def pow(x, y):  # noqa: A001 -- the reference's name
    return _map("pow", x, y)
