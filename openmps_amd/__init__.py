"""MI355X-native implementation of the OpenMps per-timestep MPS path.

The product is ``libmps_hip.so`` (C ABI in ``include/mps.h``): hand-written HIP
kernels for gfx950 driven by a thin host layer.  This package is the Python
mirror of the reference's ``Environment`` / ``Computer`` interface over that
ABI, plus the synthetic scene generator used by the benchmarks.
"""
from .computer import (Batch, Computer, ComputerException, CreateComputer, Environment, Grid,
                       GridException, MpsError, ParticleType, device_copy_probe,
                       device_math_check, make_environment_from)
from . import scenes

__all__ = ["Batch", "Computer", "ComputerException", "CreateComputer", "Environment", "Grid",
           "GridException", "MpsError", "ParticleType", "device_copy_probe", "device_math_check", "scenes",
           "make_environment"]


make_environment = make_environment_from
