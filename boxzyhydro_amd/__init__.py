"""boxzyhydro_amd -- the BOXZYHYDRO hydro step on AMD Instinct MI355X (gfx950).

Hand-written HIP kernels behind a C-ABI (include/boxzy.h, libboxzy_hip.so),
a Fortran ISO_C_BINDING driver (fortran/), and this Python binding:

* :mod:`boxzyhydro_amd.params` -- namelist values of the step
* :mod:`boxzyhydro_amd.ics`    -- set-ups (Sod tube, random box, Riemann batch)
* :mod:`boxzyhydro_amd.hydro`  -- ``HydroGrid``: the reference's subroutines on the GPU
"""
from .params import HydroParams, box_params, sod_params  # noqa: F401

__all__ = ["HydroParams", "box_params", "sod_params"]
