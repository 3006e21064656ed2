"""The C-ABI library loads and exports every function include/mps.h declares
(no compute: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mps.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(mps_\w+)\s*\(", text, flags=re.M))
    return sorted(n for n in names if not n.endswith("_fn"))


def test_header_declares_entry_points():
    names = declared()
    for must in ("mps_create", "mps_forward_time", "mps_add_particles", "mps_get_particles",
                 "mps_stage_search", "mps_set_ppe", "mps_set_solver"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from openmps_amd import _lib
    lib = _lib.lib()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # every symbol bound by the Python mirror is declared in the header
    assert set(_lib.SIGNATURES) <= set(declared())


def test_library_is_gfx950_code_object():
    from openmps_amd import _lib
    out = subprocess.run(["strings", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


def test_env_init_matches_oracle():
    """Host-side environment derivation of the library equals the oracle's."""
    import oracle
    from openmps_amd import Environment
    import cases
    for a in (cases.NUMBER_DENSITY["env"], cases.IMPLICIT["env"],
              dict(cases.NUMBER_DENSITY["env"], r_lap_by_l0=3.1)):
        e = Environment(a["max_dt"], a["courant"], a["g"], a["rho"], a["nu"], a["r_e_by_l0"],
                        a["l0"], a["min_x"], a["min_z"], a["max_x"], a["max_z"],
                        a.get("r_lap_by_l0")).as_dict()
        o = oracle.Oracle(a, 1e-10).env()
        for k in ("max_dt", "max_dx", "r_e", "r_lap", "gx", "gz", "neighbor_length", "n0",
                  "n0_lap"):
            assert e[k] == o[k], k


def test_no_device_error_without_gpu():
    """Creating a context without a GPU reports MPS_ERR_NO_DEVICE, never falls
    back to a CPU path."""
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("a GPU is present")
    from openmps_amd import Computer, Environment, MpsError
    env = Environment(0.01, 0.1, 9.8, 998.2, 1e-6, 2.1, 0.01, 0, 0, 1, 1)
    with pytest.raises(MpsError):
        Computer(1e-10, env)
