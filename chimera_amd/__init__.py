"""chimera_amd -- MI355X (gfx950) implementation of CHIMERA's SSP-RK3 stage update.

Modules:
  _lib      ctypes binding of libchimera_hip.so (include/chimera_hip.h)
  solver    resident single-grid driver (Grid)
  dropin    per-routine drop-ins with the reference's argument meaning
  problems  initial conditions (OT, rotor, KH, field loops, GRF ensemble members and
            the reference's random problem on the GPU)
  batch     batched GRF ensemble
  adjoint   hand-written adjoint behind torch.autograd
  h5io      snapshot files in the reference's HDF5 layout (libchimera_h5.so)
  cli       the reference's command line on the resident path
"""
from . import problems  # noqa: F401

__all__ = ["problems", "solver", "dropin"]


def __getattr__(name):
    # solver / dropin load the HIP library lazily (and raise if it is missing)
    if name in ("solver", "dropin", "batch", "adjoint"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
