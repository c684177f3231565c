"""pydyn_amd: the TM5 transport step of pera-noaa/pydyn, re-built for AMD Instinct MI355X.

FP64 HIP kernels for gfx950 (libtm5hip.so, C ABI in include/tm5hip.h), bit-identical to the
reference's NumPy implementation, behind the reference's operator names:

    pydyn_amd.gpu.kernels   the device operators (advect_x/y/z, convec, dynam0, ...)
    pydyn_amd.region        a global region resident on the device, run by the native driver
    pydyn_amd.meteo         the deterministic synthetic meteorology the runs are built from
    pydyn_amd.device        device arrays (torch tensors, i fastest) and the library context
    pydyn_amd.build         the hipcc build of the library
"""

__version__ = "0.1.0"
