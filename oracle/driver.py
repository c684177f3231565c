"""One global region's run: Setup_MassFlow, the two half steps of a dynamic step, and the
interval of nread seconds that Check_CFL cuts into n steps of ndyn.

The operator sequence is the zoom driver's on a tree of one region (pydyn/zoom/steps.py:259-318,
pydyn/zoom/tree.py:18): half step 1 runs "xyzvsc" then m2phlb1, half step 2 "csvzyx" then
m2phlb1.  Sources and chemistry are the forward run's callbacks (pydyn/forward.py:411-451):
`s` adds the emission of the half step at level 1 of tracer 1, `c` applies the decay of the
half step (dt = ndyn / 2) to every tracer.
"""

from . import physics
from .advection import advect_x, advect_y, advect_z, dynam0
from .cfl import check_cfl
from .convection import convec

TREF = 1  # a global region is not refined in time


class Run:
    """The state and fixed inputs of one region, built from a pydyn_amd.meteo.Meteo."""

    def __init__(self, met, limits=True):
        g = met.g
        self.met, self.g, self.limits = met, g, limits
        self.st = {k: v.copy() for k, v in met.state().items()}
        self.st["pu"], self.st["pv"] = met.fluxes(0.0)
        self.st["am"], self.st["bm"], self.st["cm"] = dynam0(self.st["pu"], self.st["pv"], g.bt, 1800.0)
        self.st["sp"], self.st["phlb"] = physics.m2phlb1(self.st["m"], g.dxyp, g.at, g.bt)
        # Kzz is a fixed profile here, so dkg is computed once from the meteo's own air mass
        self.dkg = physics.dkg(met.kvh, met.m, met.gph, g.lmax)
        self.t = 0.0
        self.log = []  # (t, ndyn, n) of every interval

    # ------------------------------------------------------------------ the operators
    def setup_massflow(self, ndyn):
        st, g = self.st, self.g
        st["am"], st["bm"], st["cm"] = dynam0(st["pu"], st["pv"], g.bt, ndyn / (2.0 * TREF))

    def _fields(self):
        st = self.st
        return st["m"], st["rm"], st["rxm"], st["rym"], st["rzm"]

    def operator(self, op, ndyn):
        st, met, g = self.st, self.met, self.g
        if op == "x":
            advect_x(*self._fields(), st["am"], self.limits)
        elif op == "y":
            advect_y(*self._fields(), st["bm"], self.limits)
        elif op == "z":
            advect_z(*self._fields(), st["cm"], self.limits)
        elif op == "v":
            convec(*self._fields(), met.entu, met.detu, met.entd, met.detd, met.cloud_top, met.cloud_lfs,
                   self.dkg, ndyn / (2.0 * TREF), g.lmax)  # fmt: skip
        elif op == "s":
            factor = physics.month2dt(ndyn, SEC_MONTH, TREF)
            x = physics.inject(met.emis, factor, XMCH4, XMCH4, 1.0)
            physics.add_2d(st["rm"], st["rzm"], x, 1, 1)
        elif op == "c":
            factor = physics.decay(met.loss, ndyn / 2.0)
            physics.apply_decay((st["rm"], st["rxm"], st["rym"], st["rzm"]), factor, range(1, met.nt + 1))
        else:
            raise ValueError(f"unknown operator {op!r}")

    def step(self, ndyn):
        """One dynamic step: "xyzvsc", m2phlb1, "csvzyx", m2phlb1."""
        g = self.g
        for half in ("xyzvsc", "csvzyx"):
            for op in half:
                self.operator(op, ndyn)
            self.st["sp"], self.st["phlb"] = physics.m2phlb1(self.st["m"], g.dxyp, g.at, g.bt)

    def interval(self, nread=3600, ndyn_max=3600):
        """Check_CFL over the coming nread seconds, then its n steps of ndyn, each with the
        fluxes of its middle (pydyn/forward.py:388-451).  Returns (ndyn, n)."""
        met, t0 = self.met, self.t

        def fluxes_at(i, ndyn):
            return met.fluxes(t0 + (i - 0.5) * ndyn)

        ndyn, n, _ = check_cfl(self.st, fluxes_at, ndyn_max, nread // ndyn_max, nread, self.g.bt)
        for s in range(n):
            self.st["pu"], self.st["pv"] = met.fluxes(t0 + s * ndyn + ndyn / 2.0)
            self.setup_massflow(ndyn)
            self.step(ndyn)
        self.t = t0 + n * ndyn
        self.log.append((t0, ndyn, n))
        return ndyn, n


SEC_MONTH = 30.0 * 86400.0
XMCH4 = 12.01115 + 1.0079 * 4.0
