"""MI355X-native relaxing-TVD split MHD step (uelipen/Heterogeneous_MHD).

The compute path is libmhd_hip.so (hand-written gfx950 HIP kernels behind the
reference's C call surface, include/mhd_hip.h).  This package is a thin
Python layer over it: ``ops`` mirrors the reference's operators over GPU
tensors, ``solver`` drives full steps, the initial condition and the FP64
diagnostics on the device.
"""
from ._lib import MHDError, lib  # noqa: F401

__all__ = ["MHDError", "lib", "ops", "solver"]


def __getattr__(name):
    if name in ("ops", "solver"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
