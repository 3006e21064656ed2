"""Host-side mirror of OpenMps's ``Environment`` and ``Computer`` over the C ABI.

Names, argument meaning and error behaviour follow the reference:
``Environment(maxDt, courant, g, rho, nu, r_eByl_0, l_0, minX, minZ, maxX, maxZ)``
(Environment.hpp:101-216), ``CreateComputer(allowableResidual, env,
positionWall, positionWallPre)`` (Computer.hpp:1800-1815), ``AddParticles``,
``ForwardTime(dt)`` / ``ForwardTime()``, ``Particles()``, ``GetEnvironment()``,
``Computer.Exception`` (CG failure) and ``Grid.Exception`` (cell overflow).
The private stages the reference's tests reach through friend classes are
exposed under the same names (``SearchNeighbor``, ``ComputeNeighborDensities``,
...).  All arithmetic runs in libmps_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional

import numpy as np

from . import _lib as L

FIELDS = ("x", "z", "u", "w", "p", "n")


class ParticleType:
    """Particle::Type, Particle.hpp:16-29."""
    IncompressibleNewton = 0
    Wall = 1
    Dummy = 2
    Disabled = 3


class MpsError(RuntimeError):
    """Any failure reported by the library."""


class ComputerException(MpsError):
    """Computer::Exception (Computer.hpp:1792-1798): the PPE solve did not converge."""


class GridException(MpsError):
    """Grid::Exception (Grid.hpp:122-129): more particles in a cell than it holds."""


class Grid:
    Exception = GridException


def _raise(rc: int, handle) -> None:
    if rc == L.MPS_OK:
        return
    msg = L.lib().mps_last_error(handle).decode() if handle else f"status {rc}"
    if rc == L.MPS_ERR_CG_NOT_CONVERGED:
        raise ComputerException(msg)
    if rc == L.MPS_ERR_CELL_OVERFLOW:
        raise GridException(msg)
    if rc == L.MPS_ERR_NO_DEVICE:
        raise MpsError("no HIP device available")
    raise MpsError(f"mps error {rc}: {msg}")


class Environment:
    """Environment (Environment.hpp:9-256) with the optional Laplacian radius extension."""

    def __init__(self, maxDt: float, courant: float, g: float, rho: float, nu: float,
                 r_eByl_0: float, l_0: float, minX: float, minZ: float, maxX: float, maxZ: float,
                 r_lapByl_0: Optional[float] = None):
        self._e = L.MpsEnv()
        _raise(L.lib().mps_env_init(ctypes.byref(self._e), maxDt, courant, g, rho, nu, r_eByl_0,
                                    l_0, minX, minZ, maxX, maxZ), None)
        if r_lapByl_0 is not None and r_lapByl_0 != r_eByl_0:
            _raise(L.lib().mps_env_set_laplacian_radius(ctypes.byref(self._e), r_lapByl_0), None)
        self._t = 0.0
        self._dt = 0.0

    @classmethod
    def _from_struct(cls, e: "L.MpsEnv", t: float, dt: float) -> "Environment":
        obj = cls.__new__(cls)
        obj._e = e
        obj._t = t
        obj._dt = dt
        return obj

    # constructor-derived constants, Environment.hpp:44-90
    MaxDt = property(lambda s: s._e.max_dt)
    MaxDx = property(lambda s: s._e.max_dx)
    L_0 = property(lambda s: s._e.l0)
    R_e = property(lambda s: s._e.r_e)
    R_lap = property(lambda s: s._e.r_lap)
    G = property(lambda s: (s._e.gx, s._e.gz))
    Rho = property(lambda s: s._e.rho)
    Nu = property(lambda s: s._e.nu)
    MinX = property(lambda s: (s._e.min_x, s._e.min_z))
    MaxX = property(lambda s: (s._e.max_x, s._e.max_z))
    NeighborLength = property(lambda s: s._e.neighbor_length)

    def N0(self) -> float:
        return self._e.n0

    def N0Lap(self) -> float:
        return self._e.n0_lap

    def T(self) -> float:
        return self._t

    def Dt(self) -> float:
        return self._dt

    def as_dict(self) -> dict:
        return {f: getattr(self._e, f) for f, _ in L.MpsEnv._fields_}


def _as_arrays(particles) -> dict:
    """Accepts a dict of arrays (x, z, u, w, p, n, type) or a sequence of
    Particle-like objects with those attributes."""
    if isinstance(particles, dict):
        out = {k: np.ascontiguousarray(particles[k], dtype=np.float64) for k in FIELDS}
        out["type"] = np.ascontiguousarray(particles["type"], dtype=np.uint8)
        return out
    rows = list(particles)
    out = {k: np.array([getattr(p, k) for p in rows], dtype=np.float64) for k in FIELDS}
    out["type"] = np.array([p.type for p in rows], dtype=np.uint8)
    return out


class Computer:
    """Computer<PW, PWP> (Computer.hpp:434-1790) on one MI355X."""

    Exception = ComputerException

    def __init__(self, allowableResidual: float, env: Environment,
                 positionWall: Optional[Callable] = None,
                 positionWallPre: Optional[Callable] = None, device: int = 0):
        self._h = None
        h = ctypes.c_void_p()
        _raise(L.lib().mps_create(ctypes.byref(env._e), allowableResidual, device, ctypes.byref(h)),
               None)
        self._h = h
        self._env = env
        self._eps = allowableResidual
        self._wall_cbs = None
        if positionWall is not None or positionWallPre is not None:
            self._install_wall_callbacks(positionWall, positionWallPre)

    def __del__(self):
        if getattr(self, "_h", None):
            L.lib().mps_destroy(self._h)
            self._h = None

    # -- wall kinematics ------------------------------------------------------
    def _install_wall_callbacks(self, pw, pwp) -> None:
        def wall(user, t, dt, count, idx, xo, zo):
            for k in range(count):
                px, pz = pw(int(idx[k]), t, dt)
                xo[k] = px
                zo[k] = pz

        def pre(user, t, dt):
            if pwp is not None:
                pwp(t, dt)

        fw = L.WALL_FN(wall) if pw is not None else ctypes.cast(None, L.WALL_FN)
        fp = L.WALL_PRE_FN(pre)
        self._wall_cbs = (fw, fp)  # keep the thunks alive
        _raise(L.lib().mps_set_wall_callbacks(self._h, fw, fp, None), self._h)

    def set_wall_motion(self, motions, group) -> None:
        """Device-side wall kinematics (mps_set_wall_motion): non-fluid
        particle i of group g = group[i] >= 1 moves to its initial position plus
        (vx t + ax sin(omega t + phase), vz t + az sin(omega t + phase)) of
        motions[g - 1] (dicts with those keys; missing keys are 0).  An empty
        `motions` switches the motion off."""
        m = (L.WallMotion * max(len(motions), 1))()
        for k, d in enumerate(motions):
            m[k] = L.WallMotion(*(float(d.get(f, 0.0)) for f, _ in L.WallMotion._fields_))
        g = np.ascontiguousarray(group, dtype=np.uint8)
        _raise(L.lib().mps_set_wall_motion(self._h, len(motions), m,
                                           g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))), self._h)

    # -- reference interface ----------------------------------------------------
    def AddParticles(self, particles) -> None:
        a = _as_arrays(particles)
        n = len(a["x"])
        _raise(L.lib().mps_add_particles(self._h, n, *[a[k].ctypes.data for k in FIELDS],
                                         a["type"].ctypes.data), self._h)

    def ForwardTime(self, dt: Optional[float] = None) -> float:
        if dt is None:
            out = ctypes.c_double()
            _raise(L.lib().mps_forward_time_cfl(self._h, ctypes.byref(out)), self._h)
            return out.value
        _raise(L.lib().mps_forward_time(self._h, dt), self._h)
        return dt

    def Particles(self) -> dict:
        return self.particles()

    def GetEnvironment(self) -> Environment:
        e = L.MpsEnv()
        _raise(L.lib().mps_get_env(self._h, ctypes.byref(e)), self._h)
        t, dt = self.time()
        return Environment._from_struct(e, t, dt)

    def DetermineDt(self) -> float:
        out = ctypes.c_double()
        _raise(L.lib().mps_determine_dt(self._h, ctypes.byref(out)), self._h)
        return out.value

    # friend-access stages (Computer.hpp:437-460)
    def SearchNeighbor(self):
        self.stage("search")

    def ComputeNeighborDensities(self):
        self.stage("density")

    def ComputeErrorCorrection(self):
        self.stage("error_correction")

    def ComputeExplicitForces(self):
        self.stage("explicit")

    def SaveX(self):
        self.stage("save_x")

    def SetPressurePoissonEquation(self):
        self.stage("ppe_assemble")

    def SolvePressurePoissonEquation(self):
        self.stage("ppe_solve")

    def ComputeImplicitForces(self):
        for s in ("ppe_assemble", "ppe_solve", "pressure_writeback", "pressure_gradient"):
            self.stage(s)

    def ModifyByPressureGradient(self):
        self.stage("pressure_gradient")

    def DynamicStabilize(self):
        self.stage("dynamic_stabilize")

    def NeighborDensityVariationSpeed(self, i: Optional[int] = None):
        v = self.density_variation()
        return v if i is None else float(v[i])

    def NeighborCount(self, i: int) -> int:
        cnt, _ = self.neighbors()
        return int(cnt[i])

    # -- generic interface (same names as the oracle wrapper) --------------------
    def size(self) -> int:
        return L.lib().mps_size(self._h)

    def add_particles(self, s: dict) -> None:
        self.AddParticles(s)

    def particles(self) -> dict:
        n = self.size()
        out = {k: np.empty(n) for k in FIELDS}
        out["type"] = np.empty(n, dtype=np.uint8)
        _raise(L.lib().mps_get_particles(self._h, *[out[k].ctypes.data for k in FIELDS],
                                         out["type"].ctypes.data), self._h)
        return out

    def set_particles(self, s: dict) -> None:
        keep = {k: np.ascontiguousarray(s[k], dtype=np.float64) for k in FIELDS if k in s}
        if "type" in s:
            keep["type"] = np.ascontiguousarray(s["type"], dtype=np.uint8)
        ptrs = [keep[k].ctypes.data if k in keep else None for k in FIELDS + ("type",)]
        _raise(L.lib().mps_set_particles(self._h, *ptrs), self._h)

    def aux(self, name: str) -> np.ndarray:
        n = self.ppe_dim() if name in ("ppe_b", "ppe_x") else self.size()
        out = np.empty(n)
        _raise(L.lib().mps_get_aux(self._h, L.AUX[name], out.ctypes.data), self._h)
        return out

    def set_aux(self, name: str, v) -> None:
        v = np.ascontiguousarray(v, dtype=np.float64)
        _raise(L.lib().mps_set_aux(self._h, L.AUX[name], v.ctypes.data), self._h)

    def set_time(self, t: float, dt: float) -> None:
        _raise(L.lib().mps_set_time(self._h, t, dt), self._h)

    def time(self):
        t, dt = ctypes.c_double(), ctypes.c_double()
        _raise(L.lib().mps_get_time(self._h, ctypes.byref(t), ctypes.byref(dt)), self._h)
        return t.value, dt.value

    def env(self) -> dict:
        e = L.MpsEnv()
        _raise(L.lib().mps_get_env(self._h, ctypes.byref(e)), self._h)
        return {f: getattr(e, f) for f, _ in L.MpsEnv._fields_}

    def grid(self) -> dict:
        out = np.zeros(4, dtype=np.int64)
        _raise(L.lib().mps_grid_shape(self._h, out.ctypes.data), self._h)
        return dict(zip(("nx", "nz", "cap", "kmax"), out.tolist()))

    def stage(self, name: str) -> None:
        _raise(getattr(L.lib(), "mps_stage_" + name)(self._h), self._h)

    def density_variation(self) -> np.ndarray:
        out = np.empty(self.size())
        _raise(L.lib().mps_density_variation(self._h, out.ctypes.data), self._h)
        return out

    def determine_dt(self) -> float:
        return self.DetermineDt()

    def forward_time(self, dt: Optional[float] = None) -> float:
        return self.ForwardTime(dt)

    def neighbors(self):
        n = self.size()
        cnt = np.empty(n, dtype=np.uint32)
        _raise(L.lib().mps_get_neighbor_counts(self._h, cnt.ctypes.data), self._h)
        ids = np.empty(int(cnt.sum(dtype=np.int64)), dtype=np.uint32)
        _raise(L.lib().mps_get_neighbors(self._h, ids.ctypes.data), self._h)
        return cnt, ids

    def ppe_dim(self) -> int:
        n = ctypes.c_size_t()
        if L.lib().mps_ppe_shape(self._h, ctypes.byref(n), None) != L.MPS_OK:
            return 0
        return n.value

    def ppe(self):
        nnz = ctypes.c_size_t()
        nn = ctypes.c_size_t()
        _raise(L.lib().mps_ppe_shape(self._h, ctypes.byref(nn), ctypes.byref(nnz)), self._h)
        n = nn.value
        rp = np.empty(n + 1, dtype=np.uint64)
        col = np.empty(nnz.value, dtype=np.uint32)
        val = np.empty(nnz.value)
        _raise(L.lib().mps_get_ppe(self._h, rp.ctypes.data, col.ctypes.data, val.ctypes.data),
               self._h)
        return rp, col, val, self.aux("ppe_b"), self.aux("ppe_x")

    def set_ppe(self, rowptr, col, val, b, x0) -> None:
        rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        _raise(L.lib().mps_set_ppe(self._h, len(b), rowptr.ctypes.data, col.ctypes.data,
                                   val.ctypes.data, b.ctypes.data, x0.ctypes.data), self._h)

    # -- solver extension and measurement ------------------------------------------
    def set_solver(self, kind: int = L.SOLVER_CG_FP64, stop_rule: int = L.STOP_INITIAL_RESIDUAL,
                   warm_start: int = L.WARM_PREVIOUS, jacobi: bool = False, tol: float = 0.0,
                   max_iter: int = 0, inner_tol: float = 1e-4, inner_max_iter: int = 0,
                   precond: int | None = None) -> None:
        """`precond` (mps_precond) overrides `jacobi` (PRECOND_JACOBI when true)."""
        if precond is None:
            precond = L.PRECOND_JACOBI if jacobi else L.PRECOND_NONE
        o = L.SolverOpts(kind, stop_rule, warm_start, int(precond), tol, max_iter, inner_tol,
                         inner_max_iter)
        _raise(L.lib().mps_set_solver(self._h, ctypes.byref(o)), self._h)

    def solve_stats(self) -> dict:
        s = L.SolveStats()
        _raise(L.lib().mps_get_solve_stats(self._h, ctypes.byref(s)), self._h)
        return {f: getattr(s, f) for f, _ in L.SolveStats._fields_}

    def cg_stats(self) -> dict:
        s = self.solve_stats()
        return {"iterations": int(s["iterations"]), "rr0": s["rr0"], "rr": s["rr"]}

    def set_storage_order(self, keep: bool) -> None:
        """keep=False (default): particles stored in cell order from each
        neighbour search on; keep=True: the current storage order is kept."""
        _raise(L.lib().mps_set_storage_order(self._h, L.ORDER_KEEP if keep else L.ORDER_CELL), self._h)

    def set_profiling(self, on: bool) -> None:
        _raise(L.lib().mps_set_profiling(self._h, int(on)), self._h)

    def reset_kernel_stats(self) -> None:
        _raise(L.lib().mps_reset_kernel_stats(self._h), self._h)

    def replay_spmv(self, kernel: str, reps: int = 50) -> float:
        """Mean ms of one launch of the last solve's SpMV ("cg_spmv32" or
        "cg_spmv"), re-launched `reps` times back to back between two HIP
        events on the context stream (mps_replay_spmv)."""
        out = ctypes.c_double()
        _raise(L.lib().mps_replay_spmv(self._h, L.KERNELS.index(kernel), int(reps), ctypes.byref(out)), self._h)
        return out.value

    def kernel_stats(self) -> dict:
        out = {}
        for k in range(len(L.KERNELS)):
            s = L.KernelStat()
            _raise(L.lib().mps_get_kernel_stat(self._h, k, ctypes.byref(s)), self._h)
            out[s.name.decode()] = {"launches": s.launches, "ms": s.total_ms,
                                    "bytes": s.total_bytes}
        return out


class Batch:
    """Independent scenes stepped together (mps_batch_*): the stages before
    and after the pressure solve per scene from `host_threads` host threads,
    the solves of the scenes on the mixed multilevel solver as grouped
    launches over all scenes (per-scene alpha, beta and stop test), one MLP
    inference over all learned scenes' rows.  Holds references to its scenes."""

    def __init__(self, scenes=(), host_threads: int = 8, device: int = 0):
        self._h = None
        h = ctypes.c_void_p()
        rc = L.lib().mps_batch_create(device, host_threads, ctypes.byref(h))
        if rc != L.MPS_OK:
            raise MpsError("no HIP device available" if rc == L.MPS_ERR_NO_DEVICE else f"mps error {rc}")
        self._h = h
        self.scenes = []
        for c in scenes:
            self.add(c)

    def __del__(self):
        if getattr(self, "_h", None):
            L.lib().mps_batch_destroy(self._h)
            self._h = None

    def _check(self, rc: int) -> None:
        if rc == L.MPS_OK:
            return
        msg = L.lib().mps_batch_last_error(self._h).decode()
        if rc == L.MPS_ERR_CG_NOT_CONVERGED:
            raise ComputerException(msg)
        if rc == L.MPS_ERR_CELL_OVERFLOW:
            raise GridException(msg)
        raise MpsError(f"mps error {rc}: {msg}")

    def add(self, c: "Computer") -> None:
        self._check(L.lib().mps_batch_add(self._h, c._h))
        self.scenes.append(c)

    def forward_time(self, dt: Optional[float] = None) -> np.ndarray:
        """ForwardTime of every scene: dt the fixed step, None each scene's
        CFL step; returns the steps taken."""
        out = np.zeros(len(self.scenes))
        self._check(L.lib().mps_batch_forward_time(self._h, dt if dt is not None else 0.0,
                                                   out.ctypes.data))
        return out

    def solve_info(self) -> dict:
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self._check(L.lib().mps_batch_solve_info(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return {"solve_ms": ms.value, "launches": n.value}


def make_environment_from(env: dict) -> Environment:
    """Environment from a dictionary with the keys of scenes.config(...)["env"]."""
    return Environment(env["max_dt"], env["courant"], env["g"], env["rho"], env["nu"],
                       env["r_e_by_l0"], env["l0"], env["min_x"], env["min_z"], env["max_x"],
                       env["max_z"], env.get("r_lap_by_l0"))


def CreateComputer(allowableResidual: float, env: Environment, positionWall=None,
                   positionWallPre=None, device: int = 0) -> Computer:
    """CreateComputer, Computer.hpp:1800-1815."""
    return Computer(allowableResidual, env, positionWall, positionWallPre, device)


def device_math_check(a, b):
    """sqrt(a) and a / b computed on the GPU by the kernels' own code path."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    s = np.empty_like(a)
    d = np.empty_like(a)
    _raise(L.lib().mps_device_math_check(len(a), a.ctypes.data, b.ctypes.data, s.ctypes.data,
                                         d.ctypes.data), None)
    return s, d


def device_copy_probe(nbytes: int) -> float:
    """Duration (ms) of a plain device-to-device copy kernel moving nbytes."""
    out = ctypes.c_double()
    _raise(L.lib().mps_device_copy_probe(nbytes, ctypes.byref(out)), None)
    return out.value
