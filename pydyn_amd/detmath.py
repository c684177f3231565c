"""Transcendental functions built from IEEE add, subtract, multiply and divide only.

The synthetic inputs have to be the same bits on every x86 host that builds them (this
container and a GPU box with another CPU).  NumPy's vectorised sin/exp/log may take different
SIMD paths on different CPUs, and glibc picks FMA or non-FMA variants at load time, so neither
is used here.  Every function below is a range reduction plus a truncated Taylor / atanh
series evaluated with NumPy's elementwise +, -, *, / (each correctly rounded, no contraction)
and exact helpers (floor, ldexp, frexp).  They are accurate to a few ulp, which is all the
generator needs; what matters is that they are deterministic.
"""

import numpy as np

__all__ = ["cos", "exp", "log", "sin"]

_PIO2_HI = 1.5707963267948966  # pi/2 rounded to double
_PIO2_LO = 6.123233995736766e-17  # pi/2 - _PIO2_HI
_LN2_HI = 0.6931471805599453
_LN2_LO = 2.3190468138462996e-17


def _poly_sin(r):
    r2 = r * r
    acc = np.zeros_like(r)
    # r - r^3/3! + r^5/5! ... up to r^19/19!, Horner from the top
    for k in range(19, 0, -2):
        fact = 1.0
        for q in range(2, k + 1):
            fact = fact * q
        term = (1.0 if ((k - 1) // 2) % 2 == 0 else -1.0) / fact
        acc = acc * r2 + term if k != 19 else np.full_like(r, term)
    return acc * r


def _poly_cos(r):
    r2 = r * r
    acc = None
    for k in range(20, -1, -2):
        fact = 1.0
        for q in range(2, k + 1):
            fact = fact * q
        term = (1.0 if (k // 2) % 2 == 0 else -1.0) / fact
        acc = np.full_like(r, term) if acc is None else acc * r2 + term
    return acc


def _quadrant(x):
    x = np.asarray(x, dtype=np.float64)
    k = np.floor(x * (2.0 / np.pi) + 0.5)
    r = (x - k * _PIO2_HI) - k * _PIO2_LO
    q = np.mod(k, 4.0)
    return r, q


def sin(x):
    r, q = _quadrant(x)
    s, c = _poly_sin(r), _poly_cos(r)
    return np.where(q == 0.0, s, np.where(q == 1.0, c, np.where(q == 2.0, -s, -c)))


def cos(x):
    r, q = _quadrant(x)
    s, c = _poly_sin(r), _poly_cos(r)
    return np.where(q == 0.0, c, np.where(q == 1.0, -s, np.where(q == 2.0, -c, s)))


def exp(x):
    x = np.asarray(x, dtype=np.float64)
    k = np.floor(x * (1.0 / _LN2_HI) + 0.5)
    r = (x - k * _LN2_HI) - k * _LN2_LO
    acc = np.ones_like(r)
    for n in range(18, 0, -1):  # 1 + r(1 + r/2(1 + r/3(...)))
        acc = 1.0 + (r / n) * acc
    return np.ldexp(acc, k.astype(np.int64))


def log(x):
    x = np.asarray(x, dtype=np.float64)
    mant, e = np.frexp(x)  # x = mant * 2**e, mant in [0.5, 1)
    small = mant < 0.7071067811865476
    mant = np.where(small, mant * 2.0, mant)
    e = np.where(small, e - 1, e).astype(np.float64)
    z = (mant - 1.0) / (mant + 1.0)
    z2 = z * z
    acc = np.zeros_like(z)
    for n in range(41, 0, -2):
        acc = acc * z2 + 1.0 / n
    return (2.0 * z) * acc + e * _LN2_HI + e * _LN2_LO
