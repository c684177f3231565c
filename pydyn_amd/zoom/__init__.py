"""The zoom tree on the device: a global region with nested children, their two-way exchange and
the tree's Check_CFL (pydyn.zoom's driver and pydyn.gpu.zoom's bodies).

    tree    Region, Zoom: the regions and what the model derives from them
    state   ZoomState: every region's halo'd arrays on the device, Fortran bounds
    lib     the device bodies (pydyn.gpu.zoom's names)
    driver  do_steps / proces_region / step, setup_massflow, check_cfl_zoom
    run     ZoomRun: a two-region tree on nested synthetic meteo, interval by interval
"""

from .state import ZoomState
from .tree import N_OPERATORS, SPLITORDER, Region, Zoom

__all__ = ["N_OPERATORS", "SPLITORDER", "Region", "Zoom", "ZoomState"]
