"""ctypes wrapper of the CPU oracle (oracle/mps_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and
by the ``cpu_baseline`` leg of ``bench.py``, always as the checker or as the
timed CPU baseline, never as the product path.

The method names mirror ``openmps_amd.Computer`` so that a parity test can drive
both objects with the same calls.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libmps_oracle.so")
_lib = None

_D = ctypes.c_double
_P = ctypes.c_void_p
_SZ = ctypes.c_size_t

ENV_FIELDS = ("max_dt", "max_dx", "l0", "r_e", "r_lap", "gx", "gz", "rho", "nu", "min_x",
              "min_z", "max_x", "max_z", "neighbor_length", "n0", "n0_lap", "courant")
AUX = {"n_without_spp": 0, "ecs": 1, "saved_x": 2, "saved_z": 3, "du_x": 4, "du_z": 5,
       "ppe_b": 6, "ppe_x": 7}


def build() -> None:
    """Compile the oracle with its committed Makefile."""
    import subprocess
    subprocess.run(["make", "-B", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_create.restype = _P
        L.oracle_create.argtypes = [_D] * 13
        L.oracle_destroy.argtypes = [_P]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_last_error.argtypes = [_P]
        L.oracle_env.argtypes = [_P, _P]
        L.oracle_grid.argtypes = [_P, _P]
        L.oracle_add_particles.argtypes = [_P, _SZ] + [_P] * 7
        L.oracle_size.restype = _SZ
        L.oracle_size.argtypes = [_P]
        L.oracle_get_particles.argtypes = [_P] + [_P] * 7
        L.oracle_set_particles.argtypes = [_P] + [_P] * 7
        L.oracle_get_aux.argtypes = [_P, ctypes.c_int, _P]
        L.oracle_set_aux.argtypes = [_P, ctypes.c_int, _P]
        L.oracle_set_time.argtypes = [_P, _D, _D]
        L.oracle_get_time.argtypes = [_P, _P]
        L.oracle_set_wall_positions.argtypes = [_P, _P, _P]
        for name in ("search", "ppe_solve"):
            f = getattr(L, "oracle_stage_" + name)
            f.restype = ctypes.c_int
            f.argtypes = [_P]
        for name in ("density", "error_correction", "explicit", "save_x", "ppe_assemble",
                     "pressure_writeback", "pressure_gradient", "dynamic_stabilize"):
            f = getattr(L, "oracle_stage_" + name)
            f.restype = None
            f.argtypes = [_P]
        L.oracle_density_variation.argtypes = [_P, _P]
        L.oracle_determine_dt.restype = _D
        L.oracle_determine_dt.argtypes = [_P]
        L.oracle_forward_time.restype = ctypes.c_int
        L.oracle_forward_time.argtypes = [_P, _D]
        L.oracle_forward_time_cfl.restype = ctypes.c_int
        L.oracle_forward_time_cfl.argtypes = [_P, _P]
        L.oracle_neighbor_counts.argtypes = [_P, _P]
        L.oracle_neighbors.argtypes = [_P, _P]
        L.oracle_ppe_dim.restype = _SZ
        L.oracle_ppe_dim.argtypes = [_P]
        L.oracle_ppe_nnz.restype = _SZ
        L.oracle_ppe_nnz.argtypes = [_P]
        L.oracle_get_ppe.argtypes = [_P, _P, _P, _P]
        L.oracle_set_ppe.argtypes = [_P, _SZ, _P, _P, _P, _P, _P]
        L.oracle_cg_stats.argtypes = [_P, _P]
        L.oracle_set_stop_rule.argtypes = [_P, ctypes.c_int, _D]
        L.oracle_set_verbose.argtypes = [_P, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleError(RuntimeError):
    pass


class CgNotConverged(OracleError):
    pass


class CellOverflow(OracleError):
    pass


class Oracle:
    """One CPU restatement of ``Computer`` (reference ``Computer.hpp:434-1790``)."""

    def __init__(self, env_args: dict, eps: float):
        a = env_args
        self._h = lib().oracle_create(a["max_dt"], a["courant"], a["g"], a["rho"], a["nu"],
                                      a["r_e_by_l0"], a["l0"], a["min_x"], a["min_z"],
                                      a["max_x"], a["max_z"], a.get("r_lap_by_l0") or 0.0, eps)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_destroy(h)
            self._h = None

    # -- environment ------------------------------------------------------
    def env(self) -> dict:
        out = np.zeros(len(ENV_FIELDS))
        lib().oracle_env(self._h, out.ctypes.data)
        return dict(zip(ENV_FIELDS, out.tolist()))

    def grid(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        lib().oracle_grid(self._h, out.ctypes.data)
        return dict(zip(("nx", "nz", "cap", "kmax"), out.tolist()))

    # -- particles --------------------------------------------------------
    def add_particles(self, s: dict) -> None:
        n = len(s["x"])
        arrs = [np.ascontiguousarray(s[k], dtype=np.float64) for k in ("x", "z", "u", "w", "p", "n")]
        t = np.ascontiguousarray(s["type"], dtype=np.uint8)
        lib().oracle_add_particles(self._h, n, *[a.ctypes.data for a in arrs], t.ctypes.data)

    def size(self) -> int:
        return lib().oracle_size(self._h)

    def particles(self) -> dict:
        n = self.size()
        out = {k: np.empty(n) for k in ("x", "z", "u", "w", "p", "n")}
        out["type"] = np.empty(n, dtype=np.uint8)
        lib().oracle_get_particles(self._h, *[out[k].ctypes.data for k in
                                              ("x", "z", "u", "w", "p", "n", "type")])
        return out

    def set_particles(self, s: dict) -> None:
        keep = {}
        for k in ("x", "z", "u", "w", "p", "n"):
            if k in s:
                keep[k] = np.ascontiguousarray(s[k], dtype=np.float64)
        if "type" in s:
            keep["type"] = np.ascontiguousarray(s["type"], dtype=np.uint8)
        lib().oracle_set_particles(self._h, *[_ptr(keep.get(k)) for k in
                                              ("x", "z", "u", "w", "p", "n", "type")])

    def aux(self, name: str) -> np.ndarray:
        n = self.size()
        out = np.empty(n)
        lib().oracle_get_aux(self._h, AUX[name], out.ctypes.data)
        return out

    def set_aux(self, name: str, v) -> None:
        v = np.ascontiguousarray(v, dtype=np.float64)
        lib().oracle_set_aux(self._h, AUX[name], v.ctypes.data)

    def set_time(self, t: float, dt: float) -> None:
        lib().oracle_set_time(self._h, t, dt)

    def time(self):
        out = np.zeros(2)
        lib().oracle_get_time(self._h, out.ctypes.data)
        return float(out[0]), float(out[1])

    def set_wall_positions(self, x, z) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        z = np.ascontiguousarray(z, dtype=np.float64)
        lib().oracle_set_wall_positions(self._h, x.ctypes.data, z.ctypes.data)

    # -- stages -------------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc == 1:
            raise CgNotConverged(lib().oracle_last_error(self._h).decode())
        if rc == 2:
            raise CellOverflow(lib().oracle_last_error(self._h).decode())
        if rc:
            raise OracleError(f"oracle status {rc}")

    def stage(self, name: str) -> None:
        f = getattr(lib(), "oracle_stage_" + name)
        rc = f(self._h)
        if rc is not None:
            self._check(rc)

    def density_variation(self) -> np.ndarray:
        out = np.empty(self.size())
        lib().oracle_density_variation(self._h, out.ctypes.data)
        return out

    def determine_dt(self) -> float:
        return lib().oracle_determine_dt(self._h)

    def forward_time(self, dt: float | None = None) -> float:
        if dt is None:
            out = ctypes.c_double(0.0)
            self._check(lib().oracle_forward_time_cfl(self._h, ctypes.byref(out)))
            return out.value
        self._check(lib().oracle_forward_time(self._h, dt))
        return dt

    # -- neighbour lists and PPE ------------------------------------------------
    def neighbors(self):
        n = self.size()
        cnt = np.empty(n, dtype=np.uint32)
        lib().oracle_neighbor_counts(self._h, cnt.ctypes.data)
        ids = np.empty(int(cnt.sum(dtype=np.int64)), dtype=np.uint32)
        lib().oracle_neighbors(self._h, ids.ctypes.data)
        return cnt, ids

    def ppe(self):
        n = lib().oracle_ppe_dim(self._h)
        nnz = lib().oracle_ppe_nnz(self._h)
        rp = np.empty(n + 1, dtype=np.uint64)
        col = np.empty(nnz, dtype=np.uint32)
        val = np.empty(nnz)
        lib().oracle_get_ppe(self._h, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
        b = np.empty(n)
        x = np.empty(n)
        lib().oracle_get_aux(self._h, AUX["ppe_b"], b.ctypes.data)
        lib().oracle_get_aux(self._h, AUX["ppe_x"], x.ctypes.data)
        return rp, col, val, b, x

    def set_ppe(self, rowptr, col, val, b, x0) -> None:
        rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        lib().oracle_set_ppe(self._h, len(b), rowptr.ctypes.data, col.ctypes.data,
                             val.ctypes.data, b.ctypes.data, x0.ctypes.data)

    def set_stop_rule(self, rule: int, tol: float) -> None:
        """0: the reference rule (relative to the warm start); 1: |b-Ax| <= tol |b|."""
        lib().oracle_set_stop_rule(self._h, rule, tol)

    def set_verbose(self, on: bool) -> None:
        lib().oracle_set_verbose(self._h, int(on))

    def cg_stats(self) -> dict:
        out = np.zeros(3)
        lib().oracle_cg_stats(self._h, out.ctypes.data)
        return {"iterations": int(out[0]), "rr0": float(out[1]), "rr": float(out[2])}
