"""MI355X-native implementation of the OpenMps per-timestep MPS path.

The product is ``libmps_hip.so`` (C ABI in ``include/mps.h``): hand-written HIP
kernels for gfx950 driven by a thin host layer.  This package is the Python
mirror of the reference's ``Environment`` / ``Computer`` interface over that
ABI, plus the synthetic scene generator used by the benchmarks.
"""
from .computer import (Computer, ComputerException, CreateComputer, Environment, Grid,
                       GridException, MpsError, ParticleType, device_math_check)
from . import scenes

__all__ = ["Computer", "ComputerException", "CreateComputer", "Environment", "Grid",
           "GridException", "MpsError", "ParticleType", "device_math_check", "scenes",
           "make_environment"]


def make_environment(env: dict) -> Environment:
    """Environment from a scene dictionary (keys of ``scenes.config(...)["env"]``)."""
    return Environment(env["max_dt"], env["courant"], env["g"], env["rho"], env["nu"],
                       env["r_e_by_l0"], env["l0"], env["min_x"], env["min_z"], env["max_x"],
                       env["max_z"], env.get("r_lap_by_l0"))
