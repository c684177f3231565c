"""The zoom driver on a device state: do_steps, Proces_Region and the two half steps of a step,
Setup_MassFlow's region loop with coarsen_region, and the tree's Check_CFL
(pydyn/zoom/steps.py, pydyn/zoom/massflow.py:62-193, pydyn/zoom/check_cfl.py).

The operator sequence is the reference's; two things differ in how it is issued, not in what
is computed:
  * A region's convection and the sources and chemistry next to it run in one column pass
    (lib.column_physics): "csv" and "vsc" always; "vsc" + "csv" with no air-mass change between
    them (region 1 across its two half steps, with the m2phlb1 between them folded in; a child
    within one Proces_Region) as one "vsccsv" pass whose convection matrix is inverted once.
  * Check_CFL does not read a verdict after each operator: a failing mass-only operator sets
    the context's rehearsal flag and every later rehearsal kernel does nothing, so the verdict
    is read once per attempt.  What the reference skips after a failure is either never run
    or overwritten by the restore that follows.
"""

from .. import _lib
from .. import device as D
from . import lib as L
from .tree import N_OPERATORS, split_orders

__all__ = ["check_cfl_zoom", "setup_massflow", "step", "x_scope", "y_scope", "z_scope"]

CFL_SPLITORDER, CFL_N_OPERATORS = "xyzzyx", 3  # advectm_cfl.F90's own order and block size
CFL_CAPS = (10, 1, 1)  # cfl_max_nloop_{x,y,z}


# ------------------------------------------------------------------ scopes (pydyn/zoom/advect.py:47-118)
def block_of(zoom, r, status, order, nops):
    """The operators of the current block of region r, and whether the parent step is in its
    second half."""
    t = zoom.ref(r, "t")
    start = (status // (nops * t)) * nops * t
    return order[start : start + nops], status % (nops * t) >= nops


def x_scope(zoom, r, blk):
    """The j range of dynamu: the whole column when y follows x in the block, else the rows
    inside the y refinement."""
    g = zoom.g(r)
    yr = 0 if r == 1 else zoom.ref(r, "y")
    seen, js, je = False, None, None
    for d in blk:
        if d == "x":
            seen = True
        elif d == "y":
            js, je = (1, g.jm) if not seen else (1 if g.touch_sp == 1 else yr + 1, g.jm if g.touch_np == 1 else g.jm - yr)
    return js, je


def y_scope(zoom, r, blk):
    """The i range of dynamv."""
    g = zoom.g(r)
    xr = 0 if r == 1 else zoom.ref(r, "x")
    seen, is_, ie = False, None, None
    for d in blk:
        if d == "x":
            is_, ie = (1, g.im) if (not seen or g.xcyc == 1) else (xr + 1, g.im - xr)
        elif d == "y":
            seen = True
    return is_, ie


def z_scope(zoom, r, blk):
    """The (i, j) scope of dynamw."""
    g = zoom.g(r)
    xr, yr = (0, 0) if r == 1 else (zoom.ref(r, "x"), zoom.ref(r, "y"))
    seen = False
    is_ = ie = js = je = None
    for d in blk:
        if d == "x":
            is_, ie = (1, g.im) if (not seen or g.xcyc == 1) else (xr + 1, g.im - xr)
        elif d == "y":
            js, je = (1, g.jm) if not seen else (1 if g.touch_sp == 1 else yr + 1, g.jm if g.touch_np == 1 else g.jm - yr)
        elif d == "z":
            seen = True
    return is_, ie, js, je


# ------------------------------------------------------------------ advection entry points
def advect_x_zoom(zoom, st, r, limits=True):
    """put_xedges, dynamu on the block's scope with the outer faces zeroed in the second half
    (pydyn/zoom/advect.py:160-176)."""
    L.put_xedges(zoom, st, r)
    blk, second = block_of(zoom, r, zoom.status[r], zoom.splitorderzoom[r], N_OPERATORS)
    js, je = x_scope(zoom, r, blk)
    with L.outer_x_faces_zeroed(zoom, st, r, second):
        if r == 1:
            L._dynamu_region1(zoom, st, js, je, limits)
        else:
            L._dynamu_child(zoom, st, r, js, je, 1, zoom.lm, limits)
            L._uncompress_xedges(zoom, st, r, js, je, 1, zoom.lm)
            L.mix_edges(zoom, st, r)


def advect_y_zoom(zoom, st, r, limits=True):
    """put_yedges, dynamv on its scope (pydyn/zoom/advect.py:179-194).  Region 1's faces outside
    its interior are zero, so the reference's rescaling of them is the identity."""
    L.put_yedges(zoom, st, r)
    blk, second = block_of(zoom, r, zoom.status[r], zoom.splitorderzoom[r], N_OPERATORS)
    is_, ie = y_scope(zoom, r, blk)
    with L.outer_y_faces_zeroed(zoom, st, r, second):
        if r == 1:
            L._dynamv_region1(zoom, st, limits)
        else:
            L._dynamv_child(zoom, st, r, is_, ie, 1, zoom.lm, limits)
            L._uncompress_yedges(zoom, st, r, is_, ie, 1, zoom.lm)
            L.mix_edges(zoom, st, r)


def advect_z_zoom(zoom, st, r, limits=True):
    """dynamw on its scope (pydyn/zoom/advect.py:197-206)."""
    blk, _ = block_of(zoom, r, zoom.status[r], zoom.splitorderzoom[r], N_OPERATORS)
    is_, ie, js, je = z_scope(zoom, r, blk)
    if r == 1:
        L._dynamw_region1(zoom, st, limits)
    else:
        L._dynamw_child(zoom, st, r, is_, ie, js, je, limits)
        L.mix_edges(zoom, st, r)


_ADVECT = {"x": advect_x_zoom, "y": advect_y_zoom, "z": advect_z_zoom}


# ------------------------------------------------------------------ the step (pydyn/zoom/steps.py)
class _Issue:
    """What one step has already issued ahead of the driver's position: done[r] is the first
    status of region r not yet run by a fused column pass; m2phlb1 is set when region 1's
    fused pass has done the m2phlb1 that ends its first half."""

    def __init__(self, n):
        self.done = [0] * (n + 1)
        self.m2phlb1 = False


def _column(zoom, st, r, phys, ndyn, issue):
    """The column pass starting at region r's current status: the longest sequence the order
    allows (see the module docstring); marks what it covered."""
    order, s = zoom.splitorderzoom[r], zoom.status[r]
    t = zoom.ref(r, "t")
    if order[s : s + 6] == "vsccsv" and (r == 1 or s // (N_OPERATORS * t) == (s + 5) // (N_OPERATORS * t)):
        ops, n = ("vscMcsv" if r == 1 else "vsccsv"), 6
        if r == 1:
            issue.m2phlb1 = True
    elif order[s : s + 3] in ("vsc", "csv"):
        ops, n = order[s : s + 3], 3
    elif order[s] == "v":
        ops, n = "v", 1
    else:
        raise NotImplementedError(f"do_steps: a lone {order[s]!r} (sources or chemistry apart from convection)")
    L.column_physics(zoom, st, r, phys.struct(r, ndyn), ops)
    issue.done[r] = s + n


def do_steps(zoom, st, r, phys, ndyn, limits, issue):
    """The next operators of region r up to a block boundary (edges put to the children unless
    the last was x) or a z right after y (pydyn/zoom/steps.py:31-69)."""
    order = zoom.splitorderzoom[r]
    for _ in range(N_OPERATORS):
        s = zoom.status[r]
        op = order[s]
        if s >= issue.done[r]:
            if op in _ADVECT:
                _ADVECT[op](zoom, st, r, limits)
            elif op in "vsc":
                _column(zoom, st, r, phys, ndyn, issue)
            else:
                raise ValueError(f"do_steps: strange value in splitorderzoom: {op!r}")
        zoom.status[r] += 1
        s_now = zoom.status[r]
        if s_now % N_OPERATORS == 0:
            if op != "x":
                L.put_xedges(zoom, st, r)
                L.put_yedges(zoom, st, r)
            return
        if op == "z" and order[s_now - 2] == "y":
            return


def proces_region(zoom, st, r, phys, ndyn, limits, issue):
    """tref_ rounds of do_steps, each followed by the children; then update_parent and m2phlb
    (m2phlb1 for region 1) (pydyn/zoom/steps.py:72-82)."""
    for _ in range(zoom.ref(r, "t")):
        do_steps(zoom, st, r, phys, ndyn, limits, issue)
        for c in zoom.children[r]:
            proces_region(zoom, st, c, phys, ndyn, limits, issue)
        if zoom.status[r] % N_OPERATORS != 0:
            do_steps(zoom, st, r, phys, ndyn, limits, issue)
    if r != 1:
        L.update_parent(zoom, st, r)
        L.m2phlb(zoom, st, r)
    elif issue.m2phlb1:
        issue.m2phlb1 = False  # done inside the fused column pass
    else:
        L.m2phlb1(zoom, st, 1)


def step(zoom, st, phys, ndyn, limits=True):
    """One ndyn step (fluxes set): every status reset, Proces_Region(1) for each half."""
    zoom.status = [0] * (zoom.nregions + 1)
    issue = _Issue(zoom.nregions)
    for _ in range(2):
        proces_region(zoom, st, 1, phys, ndyn, limits, issue)


# ------------------------------------------------------------------ Setup_MassFlow (pydyn/zoom/massflow.py)
def coarsen_region(zoom, st, r):
    p = zoom.g(r).parent
    if p <= 0:
        return
    L._coarsen_block(zoom, st, r)
    if p != 1:
        L.m2phlb(zoom, st, p)
    else:
        L.m2phlb1(zoom, st, p)
    coarsen_region(zoom, st, p)


def setup_massflow(zoom, st, ndyn, rehearsal=False):
    """dynam0 and coarsen_region region by region, then coarsen_ambm (massflow.py:181-193)."""
    for r in zoom.regions():
        L.dynam0(zoom, st, r, ndyn, rehearsal)
        coarsen_region(zoom, st, r)
    for r in zoom.regions():
        L.coarsen_ambm(zoom, st, r)


# ------------------------------------------------------------------ Check_CFL (pydyn/zoom/check_cfl.py)
class _CFLRun:
    def __init__(self, zoom, caps):
        self.caps = caps
        self.order = split_orders(zoom, CFL_SPLITORDER, CFL_N_OPERATORS)
        self.status = [0] * (zoom.nregions + 1)


def _mass_x(zoom, st, run, r, cap):
    L.put_xedges(zoom, st, r, tracers=False)
    blk, second = block_of(zoom, r, run.status[r], run.order[r], CFL_N_OPERATORS)
    js, je = x_scope(zoom, r, blk)
    with L.outer_x_faces_zeroed(zoom, st, r, second, rehearsal=True):
        if r == 1:
            L._dynamum_region1(zoom, st, js, je, cap)
        else:
            L._dynamum_child(zoom, st, r, js, je, 1, zoom.lm, cap)


def _mass_y(zoom, st, run, r, cap):
    L.put_yedges(zoom, st, r, tracers=False)
    blk, second = block_of(zoom, r, run.status[r], run.order[r], CFL_N_OPERATORS)
    is_, ie = y_scope(zoom, r, blk)
    with L.outer_y_faces_zeroed(zoom, st, r, second, rehearsal=True):
        if r == 1:
            L._dynamvm_region1(zoom, st, cap)
        else:
            L._dynamvm_child(zoom, st, r, is_, ie, 1, zoom.lm, cap)


def _mass_z(zoom, st, run, r, cap):
    blk, _ = block_of(zoom, r, run.status[r], run.order[r], CFL_N_OPERATORS)
    is_, ie, js, je = z_scope(zoom, r, blk)
    L._dynamwm(zoom, st, r, is_, ie, js, je, cap)
    if r != 1:
        L.mix_edges(zoom, st, r, tracers=False)


_MASS = {"x": _mass_x, "y": _mass_y, "z": _mass_z}


def _cfl_do_steps(zoom, st, run, r):
    for _ in range(CFL_N_OPERATORS):
        op = run.order[r][run.status[r]]
        _MASS[op](zoom, st, run, r, run.caps["xyz".index(op)])
        run.status[r] += 1
        if run.status[r] % CFL_N_OPERATORS == 0:
            return


def determine_cfl_iter(zoom, st, run, r):
    """The region's tref_ blocks, each followed by its children, then updatem_parent
    (check_cfl.py:248-266); after a failure every kernel is a no-op."""
    for _ in range(zoom.ref(r, "t")):
        _cfl_do_steps(zoom, st, run, r)
        for c in zoom.children[r]:
            determine_cfl_iter(zoom, st, run, c)
        if run.status[r] % CFL_N_OPERATORS != 0:
            _cfl_do_steps(zoom, st, run, r)
    if r != 1:
        L.update_parent(zoom, st, r, tracers=False)


def _clear_rehearsal():
    _lib.check(_lib.lib().tm5_ctx_clear(D.context().ptr, _lib.REHEARSAL), "tm5_ctx_clear")


def check_cfl_zoom(zoom, st, ndyn, n, nread, interpolate=None, max_iteration=100, caps=CFL_CAPS):
    """Check_CFL for the tree (check_cfl.py:269-290): per coming step Setup_MassFlow and the xyz
    and zyx rehearsals; m, am, bm, cm restored; on a failure the next valid ndyn (a multiple
    of 2 max tref) and the stored fluxes scaled.  Returns (ndyn, n, iterations)."""
    tref_max = max(zoom.g(r).tref for r in zoom.regions())
    ctx = D.context()
    for iteration in range(1, max_iteration + 1):
        run = _CFLRun(zoom, caps)
        _clear_rehearsal()
        L.save(st, "check_cfl", ("m", "am", "bm", "cm"))
        for i in range(1, n + 1):
            if interpolate is not None:
                interpolate(zoom, st, i, ndyn)
            setup_massflow(zoom, st, ndyn, rehearsal=True)
            determine_cfl_iter(zoom, st, run, 1)
            determine_cfl_iter(zoom, st, run, 1)
            run.status = [0] * (zoom.nregions + 1)
        L.restore(st, "check_cfl")
        flags = ctx.flags(clear=False)
        if flags & _lib.ERRORS:
            ctx.raise_errors()
        if not flags & _lib.REHEARSAL:
            return ndyn, n, iteration
        ndyn_old = ndyn
        ndyn = new_valid_timestep(ndyn, nread, nread, tref_max)
        n = nint(float(n * ndyn_old) / float(ndyn))
        L.scale(st, ("am", "bm", "cm"), float(ndyn) / float(ndyn_old))
    raise RuntimeError("check_cfl_zoom: exceeded maximum number of time steps")


def nint(x):
    """Fortran nint: half away from zero."""
    r = int(abs(x) + 0.5)
    return r if x >= 0 else -r


def new_valid_timestep(dtime, nread, cfl_outputstep, tref_max=1):
    """The largest dt < dtime dividing nread and cfl_outputstep that is a multiple of
    2 tref_max (pydyn/cfl.py:208-213)."""
    for dt in range(dtime - 1, 0, -1):
        if nread % dt == 0 and cfl_outputstep % dt == 0 and dt % (2 * tref_max) == 0:
            return dt
    raise ValueError("new_valid_timestep: no valid time step")
