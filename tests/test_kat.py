"""Known-answer tests of the reference's googletest suite, run on the CPU oracle
and (marked gpu) on the HIP library through the C ABI.

Each test cites the reference test it restates; constants and tolerances are
the reference's.
"""
import math

import numpy as np
import pytest

import cases
from openmps_amd import scenes

KINDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def _lattice_n0(r_by_l0: float) -> float:
    """Closed form of the reference density: sum over all lattice points
    0 < |(i, j)| < r of (r / |(i, j)| - 1), in l0 units (any order)."""
    R = int(math.ceil(r_by_l0)) + 1
    s = 0.0
    for i in range(-R, R + 1):
        for j in range(-R, R + 1):
            d = math.hypot(i, j)
            if 0 < d < r_by_l0:
                s += r_by_l0 / d - 1
    return s


@pytest.mark.parametrize("kind", KINDS)
def test_n0_known_answer(kind):
    """n0 = 6.539696962 at r_e = 2.1 l0 (test_ComputerNumberDensity.cpp:263-276)."""
    c = cases.make(kind, cases.NUMBER_DENSITY["env"], 1e-10)
    n0 = c.env()["n0"]
    assert abs(n0 - 6.539696962) < 1e-5
    assert abs(n0 - _lattice_n0(2.1)) < 1e-12
    # closed form at r_lap = 3.1 l0 as well
    e = dict(cases.NUMBER_DENSITY["env"], r_lap_by_l0=3.1)
    assert abs(cases.make(kind, e, 1e-10).env()["n0_lap"] - _lattice_n0(3.1)) < 1e-12


@pytest.mark.parametrize("kind", KINDS)
def test_constructor_fields(kind):
    """test_ComputerConstructor.cpp:112-181."""
    case = cases.CONSTRUCTOR
    a = case["env"]
    c = cases.lattice_case(kind, case, case["nx"], case["nz"])
    e = c.env()
    assert e["gx"] == 0.0 and e["gz"] == -a["g"]
    assert e["rho"] == a["rho"] and e["nu"] == a["nu"]
    assert e["l0"] == a["l0"] and e["r_e"] == a["r_e_by_l0"] * a["l0"]
    assert (e["min_x"], e["min_z"], e["max_x"], e["max_z"]) == (a["min_x"], a["min_z"], a["max_x"], a["max_z"])
    dx_courant = a["courant"] * a["l0"]
    assert e["max_dt"] <= a["max_dt"]
    assert e["max_dt"] <= math.sqrt(2.0 * dx_courant / a["g"])
    assert e["max_dx"] >= dx_courant
    assert e["neighbor_length"] >= a["r_e_by_l0"] * a["l0"] * (1 + a["courant"])
    assert c.size() == case["nx"] * case["nz"]
    P = c.particles()
    assert P["x"].min() >= 0 and P["x"].max() <= (case["nx"] - 1) * a["l0"]
    assert P["z"].min() >= 0 and P["z"].max() <= (case["nz"] - 1) * a["l0"]


@pytest.mark.parametrize("kind", KINDS)
def test_number_density_lattice(kind):
    """Neighbour symmetry, no self, and n at the centre of a 7x7 lattice
    (test_ComputerNumberDensity.cpp:209-276)."""
    case = cases.NUMBER_DENSITY
    c = cases.lattice_case(kind, case, case["nx"], case["nz"])
    c.stage("search")
    cnt, ids = c.neighbors()
    off = np.concatenate([[0], np.cumsum(cnt.astype(np.int64))])
    lists = [set(ids[off[i]:off[i + 1]].tolist()) for i in range(len(cnt))]
    for i, li in enumerate(lists):
        assert i not in li
        for j in li:
            assert i in lists[j]
    c.stage("density")
    centre = (case["nx"] - 1) // 2 * (case["nz"] + 1)
    assert abs(c.particles()["n"][centre] - 6.539696962) < 1e-5


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("field", ["linear", "polynomial"])
def test_density_variation_divergence(kind, field):
    """HS speed / n = -div u (test_ComputerNeighborDensityVariation.cpp:114-211)."""
    case = cases.HS
    if field == "linear":
        gx, gz = 1.0, -0.3
        u = lambda x, z: gx * x
        w = lambda x, z: gz * z
        div = lambda x, z: -(gx + gz)
    else:
        gx, gz = 1.0, -1.5
        u = lambda x, z: 0.5 * gx * x * x
        w = lambda x, z: gz * z ** 3 / 3.0
        div = lambda x, z: -(gx * x + gz * z * z)
    c = cases.lattice_case(kind, case, case["nx"], case["nz"], u=u, w=w, dt=case["dt"])
    c.stage("search")
    c.stage("density")
    speed = c.density_variation()
    n = c.particles()["n"]
    m, nx, l0 = case["margin"], case["nx"], case["env"]["l0"]
    for j in range(m, case["nz"] - m):
        for i in range(m, nx - m):
            idx = i + nx * j
            want = div(i * l0, j * l0)
            assert abs((speed[idx] / n[idx] - want) / want) < case["acc"]


def _ppe_lattice(kind, shape, u=None, w=None):
    case = cases.IMPLICIT
    c = cases.lattice_case(kind, case, shape[0], shape[1], u=u, w=w, dt=case["dt"])
    c.stage("search")
    c.stage("density")
    c.stage("ppe_assemble")
    rp, col, val, b, x = c.ppe()
    n = shape[0] * shape[1]
    return c, cases.dense(rp, col, val, n), b


@pytest.mark.parametrize("kind", KINDS)
def test_ppe_matrix_symmetry(kind):
    """test_ComputerImplicitForces.cpp:167-214."""
    _, A, _ = _ppe_lattice(kind, cases.IMPLICIT["small"])
    low = np.tril(A, -1) - np.tril(A.T, -1)
    assert math.sqrt((low ** 2).sum()) < 1e-3


@pytest.mark.parametrize("kind", KINDS)
def test_ppe_matrix_diag_identity(kind):
    """a_ii = -sum a_ij at the lattice centre (test_ComputerImplicitForces.cpp:216-260)."""
    nx, nz = cases.IMPLICIT["big"]
    _, A, _ = _ppe_lattice(kind, (nx, nz))
    i = (nx - 1) // 2 * (nz + 1)
    off = A[i].sum() - A[i, i]
    assert abs((A[i, i] + off) / A[i, i]) < 1e-3


@pytest.mark.parametrize("kind", KINDS)
def test_ppe_matrix_pattern(kind):
    """Non-zeros only at alive neighbours (test_ComputerImplicitForces.cpp:262-320)."""
    c, A, _ = _ppe_lattice(kind, cases.IMPLICIT["small"])
    cnt, ids = c.neighbors()
    off = np.concatenate([[0], np.cumsum(cnt.astype(np.int64))])
    n = A.shape[0]
    for i in range(n):
        nb = set(ids[off[i]:off[i + 1]].tolist())
        for j in range(n):
            if j != i and j not in nb:
                assert abs(A[i, j]) < 1e-3


@pytest.mark.parametrize("kind", KINDS)
def test_ppe_rhs_value(kind):
    """b = -rho/(dt n0) (-div u n) (test_ComputerImplicitForces.cpp:323-375)."""
    case = cases.IMPLICIT
    gx, gz = 1.0, -0.3
    nx, nz = case["big"]
    c, _, b = _ppe_lattice(kind, (nx, nz), u=lambda x, z: gx * x, w=lambda x, z: gz * z)
    e = c.env()
    n = c.particles()["n"]
    m = case["margin"]
    for iz in range(m, nz - m):
        for ix in range(m, nx - m):
            i = ix + nx * iz
            want = -e["rho"] / (case["dt"] * e["n0"]) * (-(gx + gz) * n[i])
            assert abs((b[i] - want) / want) < 1e-1


@pytest.mark.parametrize("kind", KINDS)
def test_ppe_matrix_value(kind):
    """A_ij = 3 r_e / (n0 r^3) (test_ComputerImplicitForces.cpp:378-452)."""
    case = cases.IMPLICIT
    nx, nz = case["big"]
    c, A, _ = _ppe_lattice(kind, (nx, nz), u=lambda x, z: 1.0 * x, w=lambda x, z: -0.3 * z)
    e = c.env()
    m = case["margin"]
    idx = [ix + nx * iz for iz in range(m, nz - m) for ix in range(m, nx - m)]
    checked = 0
    for a in idx:
        for bb in idx:
            if a >= bb or A[a, bb] == 0.0:
                continue
            r = math.hypot((a % nx - bb % nx) * e["l0"], (a // nx - bb // nx) * e["l0"])
            want = 3 * e["r_e"] / e["n0"] / r ** 3
            assert abs((A[a, bb] - want) / A[a, bb]) < 1e-3
            checked += 1
    assert checked > 0


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("name", ["identity", "indefinite4", "ode1d"])
def test_conjugate_gradient(kind, name):
    """test_ComputerConjugateGradient.cpp:115-229 (eps = 1e-7, accuracy 1e-3)."""
    case = cases.CG
    A, b, want = cases.cg_systems()[name]
    c = cases.make(kind, case["env"], case["eps"])
    rp, col, val = cases.csr_from_dense(A)
    c.set_ppe(rp, col, val, b, np.zeros(len(b)))
    c.stage("ppe_solve")
    x = c.ppe()[4]
    if name == "identity":
        assert np.abs(x - want).sum() / len(b) < case["acc"]
    elif name == "indefinite4":
        assert np.all(np.abs(x - want) < case["acc"])
    else:
        assert np.abs((x - want) / want).sum() / len(b) < case["acc"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("field", ["polynomial", "trigonometric"])
def test_pressure_gradient(kind, field):
    """du / (-dt/rho) = grad p (test_ComputerPressureGradient.cpp:117-253)."""
    case = cases.GRADIENT
    r_by = case["env"]["r_e_by_l0"]
    if field == "polynomial":
        dl = r_by * 10.0
        gpx, gpz = 1.0 / (dl * dl), -1.0 / (dl * dl)
        p = lambda x, z: 0.5 * x * x * gpx + z ** 3 / 3.0 * gpz
    else:
        k = 2.0 * math.pi / (10.0 * r_by)
        gpx, gpz = 1.0, -1.0
        p = lambda x, z: gpx * np.cos(k * x) + gpz * np.sin(k * z)
    nx, nz = case["nx"], case["nz"]
    c = cases.lattice_case(kind, case, nx, nz, p=p, dt=case["dt"])
    c.stage("search")
    c.stage("density")
    c.stage("pressure_gradient")
    P = c.particles()
    e = c.env()
    dt = case["dt"]
    pref = -dt / e["rho"]
    l0 = e["l0"]
    m = case["margin"]
    for j in range(m, nz - m):
        for i in range(m, nx - m):
            idx = i + nx * j
            x0, z0 = i * l0, j * l0
            xi, zi = P["x"][idx], P["z"][idx]
            if field == "polynomial":
                dpx, dpz = gpx * x0, gpz * z0 * z0
            else:
                dpx, dpz = -gpx * math.sin(k * xi) * k, gpz * math.cos(k * zi) * k
            assert abs((P["u"][idx] / pref - dpx) / dpx) < case["acc"]
            assert abs((P["w"][idx] / pref - dpz) / dpz) < case["acc"]
            assert abs(((xi - x0) - pref * dpx * dt) / (pref * dpx * dt)) < case["acc"]
            assert abs(((zi - z0) - pref * dpz * dt) / (pref * dpz * dt)) < case["acc"]
