"""Fixtures of the reference's unit tests, shared by the oracle and GPU tests.

Each entry restates the constants of one googletest fixture in
src/OpenMps/test/ (file cited) so that the same known-answer checks can run on
the CPU oracle and on the HIP library.  ``make(kind, ...)`` builds either
implementation behind the same method names.
"""
from __future__ import annotations

import numpy as np

import oracle
from openmps_amd import scenes


def env(max_dt, courant, g, rho, nu, r_e, l0, min_x, min_z, max_x, max_z, r_lap=None) -> dict:
    return dict(max_dt=max_dt, courant=courant, g=g, rho=rho, nu=nu, r_e_by_l0=r_e, l0=l0,
                min_x=min_x, min_z=min_z, max_x=max_x, max_z=max_z,
                r_lap_by_l0=(r_e if r_lap is None else r_lap))


# test_ComputerNumberDensity.cpp:7-32
NUMBER_DENSITY = dict(env=env(1 / 100, 0.1, 9.8, 998.2, 1.004e-06, 2.1, 0.1, -0.004, -0.004, 0.053, 0.1),
                      eps=1e-10, nx=7, nz=7)
# test_ComputerConstructor.cpp (constants at the top of the file)
CONSTRUCTOR = dict(env=env(1 / 100, 0.1, 9.8, 998.2, 1.004e-06, 2.4, 0.001, -0.004, -0.004, 0.053, 0.1),
                   eps=1e-10, nx=7, nz=9)
# test_ComputerImplicitForces.cpp:7-42
IMPLICIT = dict(env=env(1 / 100, 0.1, 9.8, 998.2, 1.004e-06, 5.0, 1.0, -30.0, -30.0, 30.0, 30.0),
                eps=1e-10, dt=1 / 100, small=(6, 6), big=(15, 15), margin=6)
# test_ComputerNeighborDensityVariation.cpp (constants at the top of the file)
HS = dict(env=env(1 / 100, 0.1, 0.0, 998.2, 1.004e-06, 5.0, 1.0, -50.0, -50.0, 50.0, 50.0),
          eps=1e-10, dt=1 / 100, nx=35, nz=35, margin=15, acc=0.10)
# test_ComputerPressureGradient.cpp (constants at the top of the file)
GRADIENT = dict(env=env(1 / 100, 0.1, 9.8, 998.2, 1.004e-06, 1.5, 1.0, -50.0, -50.0, 50.0, 50.0),
                eps=1e-10, dt=1 / 100, nx=20, nz=20, margin=5, acc=0.05)
# test_ComputerConjugateGradient.cpp (constants at the top of the file)
CG = dict(env=env(1 / 100, 0.1, 9.8, 998.2, 1.004e-06, 1.5, 0.01, -0.01, -0.01, 0.01, 0.01),
          eps=1e-7, dt=1 / 100, acc=1e-3)


def make(kind: str, env_args: dict, eps: float):
    """kind 'oracle' -> CPU oracle; 'gpu' -> openmps_amd.Computer."""
    if kind == "oracle":
        return oracle.Oracle(env_args, eps)
    from openmps_amd import Computer, make_environment
    return Computer(eps, make_environment(env_args))


def lattice_case(kind: str, case: dict, nx: int, nz: int, u=None, w=None, p=None, dt=None):
    c = make(kind, case["env"], case["eps"])
    if dt is not None:
        c.set_time(dt, dt)  # Dt() = dt_step; SetNextT()
    c.add_particles(scenes.lattice(nx, nz, case["env"]["l0"], u=u, w=w, p=p))
    return c


def dense(rp, col, val, n):
    A = np.zeros((n, n))
    for i in range(n):
        for k in range(int(rp[i]), int(rp[i + 1])):
            A[i, col[k]] = val[k]
    return A


def csr_from_dense(A):
    n = A.shape[0]
    rp, col, val = [0], [], []
    for i in range(n):
        for j in range(n):
            if A[i, j] != 0.0:
                col.append(j)
                val.append(A[i, j])
        rp.append(len(col))
    return (np.array(rp, dtype=np.uint64), np.array(col, dtype=np.uint32),
            np.array(val, dtype=np.float64))


def cg_systems():
    """The three systems of test_ComputerConjugateGradient.cpp:115-229 with
    their known solutions."""
    out = {}
    n = 10  # SolveIdentityMatrix, :115-141
    out["identity"] = (np.eye(n), np.arange(n, dtype=float), np.arange(n, dtype=float))
    A = np.array([[1.0, 1, 1, -1], [1, 1, -1, 1], [1, -1, 1, 1], [-1, 1, 1, 1]])  # :143-178
    out["indefinite4"] = (A, 4 * np.array([1.0, 2, 3, 4]), np.array([2.0, 4, 6, 8]))
    nx = 10  # Solve1dODE, :180-229 (matrix as the test fills it, column-wise)
    dx = 1.0 / (nx + 1)
    dx2 = dx * dx
    a, b = 0.0, 1.0
    A = np.zeros((nx, nx))
    rhs = np.zeros(nx)
    for j in range(1, nx - 1):
        A[j - 1, j] = 1.0 / dx2
        A[j, j] = -2.0 / dx2
        A[j + 1, j] = 1.0 / dx2
        rhs[j] = dx * (j + 1)
    A[0, 0] = -2.0 / dx2
    A[1, 0] = 1.0 / dx2
    A[nx - 1, nx - 1] = -2.0 / dx2
    A[nx - 2, nx - 1] = 1.0 / dx2
    rhs[0] = -1.0 / dx2 * a + dx
    rhs[nx - 1] = -1.0 / dx2 * b + (1.0 - dx)
    xs = dx * (np.arange(nx) + 1)
    out["ode1d"] = (A, rhs, xs ** 3 / 6 + (b - a - 1.0 / 6.0) * xs + a)
    return out


def brute_neighbors(s: dict, nl: float, env_args: dict, cap_grid=None):
    """Neighbour lists by the documented rule, written independently of the
    oracle: cells floor((x - min)/NL) on an (nx, nz) grid, the 9 cells in
    x-major order, ascending index inside a cell, keep j != i with
    sqrt(|x_i - x_j|^2) < NL (Grid.hpp:250-254, :393-401; Computer.hpp:720-755)."""
    x, z, t = s["x"], s["z"], s["type"]
    nx, nz = cap_grid
    bx = np.floor((x - env_args["min_x"]) / nl)
    bz = np.floor((z - env_args["min_z"]) / nl)
    ok = (t != 3) & (bx >= 0) & (bx < nx) & (bz >= 0) & (bz < nz)
    cells = {}
    for i in np.nonzero(ok)[0]:
        cells.setdefault((int(bx[i]), int(bz[i])), []).append(int(i))
    lists = []
    for i in range(len(x)):
        if not ok[i]:
            lists.append(None)
            continue
        out = []
        for ox in (-1, 0, 1):
            for oz in (-1, 0, 1):
                for j in cells.get((int(bx[i]) + ox, int(bz[i]) + oz), []):
                    if j == i:
                        continue
                    rx, rz = x[i] - x[j], z[i] - z[j]
                    if np.sqrt(rx * rx + rz * rz) < nl:
                        out.append(j)
        lists.append(out)
    return lists, ok
