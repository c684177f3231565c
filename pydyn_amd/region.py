"""One global region resident on the device, run interval by interval by the library's native
driver (tm5_region_interval): Check_CFL, then n steps of ndyn, each Setup_MassFlow + the
operators "xyzvsc" m2phlb1 "csvzyx" m2phlb1 with sources and chemistry.

The state stays in HBM from construction to the end of the run; only the (ndyn, n) of each
interval comes back to the host (the Check_CFL verdict is read once per attempt).
"""

import ctypes

import numpy as np

from . import _lib
from . import device as D
from .gpu import kernels as K
from .meteo import MOD_AMPLITUDE, MOD_PERIOD, SEC_MONTH, XMCH4

STATE = ("m", "rm", "rxm", "rym", "rzm", "am", "bm", "cm", "sp", "phlb")


class Region:
    """The device state and fixed inputs of a global region built from a meteo.Meteo."""

    def __init__(self, met, limits=True):
        g = met.g
        self.met, self.g, self.limits = met, g, limits
        init = met.state()
        up = D.upload
        self.m = up(init["m"])
        self.rm, self.rxm, self.rym, self.rzm = (up(init[k]) for k in ("rm", "rxm", "rym", "rzm"))
        self.pu0, self.pv0 = up(met.pu0), up(met.pv0)
        self.am, self.bm, self.cm = D.empty((g.im + 1, g.jm, g.lm)), D.empty((g.im, g.jm + 1, g.lm)), D.empty((g.im, g.jm, g.lm + 1))
        self.sp, self.phlb = D.empty((g.im, g.jm)), D.empty((g.im, g.jm, g.lm + 1))
        self.at, self.bt, self.dxyp = (up(np.asarray(a, dtype=np.float64)) for a in (g.at, g.bt, g.dxyp))
        self.entu, self.detu, self.entd, self.detd = (up(getattr(met, k)) for k in ("entu", "detu", "entd", "detd"))
        torch = __import__("torch")
        self.cloud_top = up(met.cloud_top.astype(np.int32)).to(torch.int32)
        self.cloud_lfs = up(met.cloud_lfs.astype(np.int32)).to(torch.int32)
        self.emis, self.loss = up(met.emis), up(met.loss)
        self.kvh, self.gph, self.m_meteo = up(met.kvh), up(met.gph), up(met.m)
        # the initial fluxes (dynam0 at 1800 s) and surface pressure, as the oracle's run starts
        K.fluxes_from_meteo(self.pu0, self.pv0, self.bt, 1800.0, self.am, self.bm, self.cm, met.amplitude(0.0))
        K.m2phlb1(self.m, self.dxyp, self.at, self.bt, self.sp, self.phlb)
        self.dkg = K.dkg(self.kvh, self.m_meteo, self.gph, g.lmax)
        self.t = 0.0
        self.log = []
        self._desc = self._describe()
        ptr = ctypes.c_void_p()
        _lib.check(_lib.lib().tm5_region_create(D.context().ptr, ctypes.byref(self._desc), ctypes.byref(ptr)),
                   "tm5_region_create")  # fmt: skip
        self._ptr = ptr

    def _describe(self):
        g, met = self.g, self.met
        d = _lib.RegionDesc()
        d.im, d.jm, d.lm, d.lmax, d.limits = g.im, g.jm, g.lm, g.lmax, int(self.limits)
        for name in ("m", "am", "bm", "cm", "pu0", "pv0", "sp", "phlb", "entu", "detu", "entd", "detd", "dkg",
                     "emis", "loss"):  # fmt: skip
            setattr(d, name, D.view(getattr(self, name)))
        d.q = D.tracers(self.rm, self.rxm, self.rym, self.rzm)
        d.cloud_top, d.cloud_lfs = D.iview(self.cloud_top), D.iview(self.cloud_lfs)
        d.at, d.bt, d.dxyp = self.at.data_ptr(), self.bt.data_ptr(), self.dxyp.data_ptr()
        d.mod_period, d.mod_amplitude = MOD_PERIOD, MOD_AMPLITUDE
        d.sec_month, d.mol_ratio = SEC_MONTH, XMCH4 / XMCH4
        return d

    def interval(self, nread=3600, ndyn_max=3600):
        """One interval of nread seconds (Check_CFL + its steps); returns (ndyn, n)."""
        return self.run(1, nread, ndyn_max)[0]

    def run(self, count, nread=3600, ndyn_max=3600):
        """count intervals of nread seconds, queued with one read of the device per batch
        (tm5_region_run); returns [(ndyn, n)] per interval."""
        ndyn, n = (ctypes.c_int * max(count, 1))(), (ctypes.c_int * max(count, 1))()
        rc = _lib.lib().tm5_region_run(self._ptr, float(self.t), int(nread), int(ndyn_max), int(count), ndyn, n)
        if rc == _lib.EFLAGS:
            D.context().raise_errors()
        _lib.check(rc, "tm5_region_run")
        out = [(ndyn[k], n[k]) for k in range(count)]
        for nd, nn in out:
            self.t = self.t + nn * nd
            self.log.append((nd, nn))
        return out

    def check(self):
        """Raise the reference's exception for any error the device recorded."""
        D.context().raise_errors()

    def state(self):
        """The prognostic state and the step's fluxes as NumPy arrays in the pydyn layout."""
        self.check()
        return {k: D.download(getattr(self, k)) for k in STATE}

    def close(self):
        if self._ptr:
            _lib.lib().tm5_region_destroy(self._ptr)
            self._ptr = None
