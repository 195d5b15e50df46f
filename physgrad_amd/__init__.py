"""physgrad_amd: MI355X-native MLS-MPM particle-grid transfer loop.

The compute lives in libphysgrad_mpm.so (hand-written HIP for gfx950 behind the
C-ABI of include/physgrad_mpm.h).  This package holds the build step, the
ctypes binding, the solver wrappers that mirror the reference interfaces and
the torch.autograd.Function of the adjoint.
"""
from ._lib import load, PgError  # noqa: F401
from . import _lib  # noqa: F401

__all__ = ["load", "PgError"]


def __getattr__(name):
    # solver wrappers need torch; import lazily so the build step does not
    if name in ("MPMContext", "MPMSolver", "SimulationParams", "make_params", "default_params"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)
