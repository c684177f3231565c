"""The adjoint of a window of the single-region transport on the device, behind a
torch.autograd.Function.

Given the air-mass trajectory (which does not depend on the tracers: every sub-cycle count comes
from mass and fluxes alone), each operator maps the tracers (rm, rxm, rym, rzm) linearly except
for the slope limiter clip(s, -|rm|, |rm|), which is piecewise linear.  The controls are the
initial tracers and the emission field (the injection is affine in it).

Forward: the window's intervals run through the library's native driver
(tm5_region_interval, as pydyn_amd.region.Region), keeping a checkpoint of the state at the
start of each interval and the (ndyn, n) Check_CFL chose.  Backward, interval by interval from
the last: the checkpoint is restored and the interval's steps replayed operator by operator
(tm5_region_ops: 'D' Setup_MassFlow, 'x' 'y' 'z', 'V' the fused middle, 'M' m2phlb1), with
each advection operator's inputs (m, its faces, the tracers) and the middle's m kept; then the
operators' adjoints run in reverse on the lambdas: tm5_advect_adjoint (one 64-lane workgroup per
line, recomputing the sub-steps' limiter branches from the operator's input) and
tm5_column_adjoint ("vscMcsv" with C^T, the emission's lambda accumulated).

The forward operators follow pydyn/kernels.py:129-457, pydyn/emis.py:52-68 and
pydyn/chem.py:168-178; the reference has no adjoint.
"""

import ctypes

import numpy as np

from . import _lib
from . import device as D
from .meteo import SEC_MONTH, XMCH4
from .region import STATE, Region

__all__ = ["Transport", "Window", "advect_adjoint", "loss", "samples", "transport"]

TRACERS = ("rm", "rxm", "rym", "rzm")
AXES = {"x": (0, "am", 50), "y": (1, "bm", 10), "z": (2, "cm", 15)}  # axis, faces, nloop cap


def advect_adjoint(op, m, flux, q, lam, limits=True, work=None):
    """The adjoint of advection operator op ('x' dynamu, 'y' dynamv, 'z' dynamw) at the forward
    inputs m (im, jm, lm), flux (its faces) and q (rm, rxm, rym, rzm; (im, jm, lm, nt)): lam,
    the lambdas after the operator, are overwritten with those before it.  work: device scratch
    of (im+1)(jm+1)(lm+1) + im jm lm doubles (allocated when None)."""
    axis, _, nmax = AXES[op]
    im, jm, lm = m.shape
    if work is None:
        work = D.empty(((im + 1) * (jm + 1) * (lm + 1) + im * jm * lm,))
    rc = _lib.lib().tm5_advect_adjoint(D.context().ptr, axis, im, jm, lm, ctypes.byref(D.view(m)),
                                       ctypes.byref(D.tracers(*q)), ctypes.byref(D.view(flux)),
                                       ctypes.byref(D.tracers(*lam)), int(limits), nmax,
                                       ctypes.c_void_p(work.data_ptr()))  # fmt: skip
    _lib.check(rc, "tm5_advect_adjoint")


class Window:
    """A window of `intervals` intervals of nread seconds of a global region built from a
    meteo.Meteo, with its forward (checkpointed) and its adjoint on the device."""

    def __init__(self, met, intervals, limits=True, nread=3600, ndyn_max=3600):
        self.met, self.g = met, met.g
        self.intervals, self.nread, self.ndyn_max = intervals, nread, ndyn_max
        self.region = Region(met, limits)
        self.limits = limits
        r = self.region
        self.start = {k: getattr(r, k).clone() for k in STATE}  # the state Region built
        g = self.g
        self.work = D.empty(((g.im + 1) * (g.jm + 1) * (g.lm + 1) + g.im * g.jm * g.lm,))
        self.checkpoints = []
        self.m_final = None
        self.generation = 0  # forwards run: a backward replays the trajectory of its own forward
        self._arena = []  # per step of an interval: its records' buffers, reused interval after interval

    # ------------------------------------------------------------------ forward
    def forward(self, init, emis):
        """Run the window from the initial tracers init (rm, rxm, rym, rzm; (im, jm, lm, nt)
        float64 device tensors) with emission field emis (im, jm); returns the final tracers."""
        r = self.region
        for k in STATE:
            getattr(r, k).copy_(self.start[k])
        for k, x in zip(TRACERS, init):
            getattr(r, k).copy_(x)
        r.emis.copy_(emis)
        r.t, r.log = 0.0, []
        self.checkpoints = []  # a new list: a graph that kept the last one keeps its trajectory
        self.generation += 1
        for _ in range(self.intervals):
            saved = {k: getattr(r, k).clone() for k in ("m", *TRACERS)}
            t0 = r.t
            ndyn, n = r.interval(self.nread, self.ndyn_max)
            self.checkpoints.append((t0, ndyn, n, saved))
        r.check()
        self.m_final = r.m.clone()
        return [getattr(r, k).clone() for k in TRACERS]

    # ------------------------------------------------------------------ backward
    def _ops(self, t, ndyn, ops):
        _lib.check(_lib.lib().tm5_region_ops(self.region._ptr, float(t), int(ndyn), ops.encode()), "tm5_region_ops")

    def _buffers(self, s):
        """The record buffers of step s of an interval (made on first use, then reused): per
        advection operator of "xyzVzyx" its m, faces and tracers, for V its m."""
        r = self.region
        while len(self._arena) <= s:
            step = []
            for op in "xyzVzyx":
                m = D.empty(tuple(r.m.shape))
                if op == "V":
                    step.append((m, None, None))
                else:
                    flux = getattr(r, AXES[op][1])
                    step.append((m, D.empty(tuple(flux.shape)), [D.empty(tuple(r.rm.shape)) for _ in TRACERS]))
            self._arena.append(step)
        return self._arena[s]

    def _replay(self, k, checkpoints=None):
        """Interval k of a forward's checkpoints (default the last forward's) again: the records
        the adjoint needs, in forward order, copied into the window's buffers (one multi-tensor
        copy per record)."""
        torch = __import__("torch")
        r = self.region
        t0, ndyn, n, saved = (self.checkpoints if checkpoints is None else checkpoints)[k]
        for name, x in saved.items():
            getattr(r, name).copy_(x)
        records = []
        for s in range(n):
            bufs = self._buffers(s)
            self._ops((t0 + float(s * ndyn)) + ndyn / 2.0, ndyn, "D")
            for op, (m, flux, q) in zip("xyzVzyx", bufs):
                if op == "V":
                    m.copy_(r.m)
                    records.append(("V", ndyn, m, None, None))
                else:
                    src = [r.m, getattr(r, AXES[op][1]), *(getattr(r, name) for name in TRACERS)]
                    torch._foreach_copy_([m, flux, *q], src)
                    records.append((op, ndyn, m, flux, q))
                self._ops(0.0, ndyn, op)
            self._ops(0.0, ndyn, "M")
        r.check()
        return records

    def _physics(self, ndyn):
        """tm5_zphysics of the region's column inputs for a step of ndyn."""
        r = self.region
        p = _lib.ZPhysics()
        for k in ("entu", "detu", "entd", "detd", "dkg", "emis", "loss"):
            setattr(p, k, D.view(getattr(r, k)))
        p.cloud_top, p.cloud_lfs = D.iview(r.cloud_top), D.iview(r.cloud_lfs)
        p.lmax = self.g.lmax
        half = float(ndyn) / 2.0
        p.dt_conv = half
        p.factor = half / SEC_MONTH  # month2dt (pydyn/emis.py:65-68)
        p.ratio, p.efrac = XMCH4 / XMCH4, 1.0
        p.level, p.tracer = 1, 1
        p.escale = None
        p.loss_stride = r.loss.stride(3)
        p.dt_chem = half
        p.at, p.bt = r.at.data_ptr(), r.bt.data_ptr()
        return p

    def advect_adjoint(self, op, m, flux, q, lam):
        """lam (after the operator) <- the lambdas before it; m, flux, q the operator's inputs."""
        advect_adjoint(op, m, flux, q, lam, self.limits, self.work)

    def column_adjoint(self, m, ndyn, lam, lam_emis):
        g = self.g
        p = self._physics(ndyn)
        rc = _lib.lib().tm5_column_adjoint(D.context().ptr, g.im, g.jm, g.lm, ctypes.byref(D.view(m)),
                                           ctypes.byref(D.tracers(*lam)), ctypes.byref(p), b"vscMcsv",
                                           ctypes.byref(D.view(lam_emis)))  # fmt: skip
        _lib.check(rc, "tm5_column_adjoint")

    def backward(self, lam_final, checkpoints=None):
        """(lambdas of the initial tracers, lambda of the emission) from those of the final
        tracers (four (im, jm, lm, nt) device tensors, consumed), along the trajectory of
        `checkpoints` (a forward's; default the last forward's)."""
        checkpoints = self.checkpoints if checkpoints is None else checkpoints
        lam = [D.empty(tuple(x.shape)).copy_(x) for x in lam_final]  # i fastest, one stride set
        lam_emis = D.zeros(tuple(self.region.emis.shape))
        for k in reversed(range(len(checkpoints))):
            for op, ndyn, m, flux, q in reversed(self._replay(k, checkpoints)):
                if op == "V":
                    self.column_adjoint(m, ndyn, lam, lam_emis)
                else:
                    self.advect_adjoint(op, m, flux, q, lam)
        D.context().raise_errors()
        return lam, lam_emis


class Transport(__import__("torch").autograd.Function):
    """final tracers = transport(initial tracers, emission) over a Window, differentiable in
    both; the backward is the device adjoint."""

    @staticmethod
    def forward(ctx, rm, rxm, rym, rzm, emis, window):
        ctx.window = window
        out = window.forward((rm, rxm, rym, rzm), emis)
        # this forward's trajectory: a later forward on the same window makes new checkpoints
        # (the air mass, m_final, does not depend on the controls and is the same for every forward)
        ctx.checkpoints = window.checkpoints
        ctx.generation = window.generation
        return tuple(out)

    @staticmethod
    def backward(ctx, *grads):
        torch = __import__("torch")
        w = ctx.window
        shape = tuple(w.region.rm.shape)
        lam = [torch.zeros(shape, dtype=torch.float64, device="cuda") if gr is None else gr for gr in grads]
        lam_init, lam_emis = w.backward(lam, ctx.checkpoints)
        return (*lam_init, lam_emis, None)


def transport(window, rm, rxm, rym, rzm, emis):
    return Transport.apply(rm, rxm, rym, rzm, emis, window)


# ------------------------------------------------------------------ the loss
def samples(g, n=500, seed=17):
    """n seeded sample columns (i, j) and observations 1.85e-6 +- 1e-8 (NumPy arrays)."""
    rng = np.random.default_rng(seed)
    i = rng.integers(0, g.im, n)
    j = rng.integers(0, g.jm, n)
    obs = 1.85e-6 + 1e-8 * rng.standard_normal(n)
    return i, j, obs


def loss(rm, m, i, j, obs):
    """Sum over the samples of (column-average tracer-1 mixing ratio - obs)^2, differentiable
    in rm (torch tensors on the device; i, j, obs from samples)."""
    torch = __import__("torch")
    ii = torch.as_tensor(i, device=rm.device)
    jj = torch.as_tensor(j, device=rm.device)
    ob = torch.as_tensor(obs, dtype=torch.float64, device=rm.device)
    q = rm[ii, jj, :, 0].sum(dim=1) / m[ii, jj, :].sum(dim=1)
    return ((q - ob) ** 2).sum()
