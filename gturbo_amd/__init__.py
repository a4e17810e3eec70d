"""gturbo_amd -- MI355X (gfx950) implementation of Gturbo's pseudospectral
time step behind the reference's subroutine interface.

* ``include/gturbo.h`` / ``libgturbo_hip.so``: the C ABI (HIP kernels).
* :class:`Solver`: the same entry points from Python.
* ``fortran/``: the Fortran driver and the drop-in subroutines.
"""
from .params import (InitialSpectrum, Params, Workload, ensemble_members,  # noqa: F401
                     kept_half_width, workloads)

__all__ = ["Params", "InitialSpectrum", "Workload", "workloads", "ensemble_members",
           "kept_half_width", "Solver"]


def __getattr__(name):
    if name == "Solver":
        from .solver import Solver
        return Solver
    raise AttributeError(name)
