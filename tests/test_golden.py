"""The oracle (CPU) and the HIP path (gpu) against the committed golden
fixtures of tests/golden/ (written by tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

import cases
from openmps_amd import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KINDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def load(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


@pytest.mark.parametrize("kind", KINDS)
def test_lattice_density_golden(kind):
    g = load("lattice_density")
    c = cases.lattice_case(kind, cases.NUMBER_DENSITY, 7, 7)
    c.stage("search")
    cnt, ids = c.neighbors()
    assert np.array_equal(cnt, g["count"]) and np.array_equal(ids, g["ids"])
    c.stage("density")
    assert np.array_equal(c.particles()["n"], g["n"])
    assert np.array_equal(c.aux("n_without_spp"), g["nws"])


@pytest.mark.parametrize("kind", KINDS)
def test_ppe_lattices_golden(kind):
    g = load("ppe_lattices")
    for name, shape, vel in (("small", (6, 6), False), ("big", (15, 15), True)):
        case = cases.IMPLICIT
        kw = dict(u=lambda x, z: 1.0 * x, w=lambda x, z: -0.3 * z) if vel else {}
        c = cases.lattice_case(kind, case, shape[0], shape[1], dt=case["dt"], **kw)
        for st in ("search", "density", "ppe_assemble"):
            c.stage(st)
        rp, col, val, b, _ = c.ppe()
        for got, key in ((rp, "rowptr"), (col, "col"), (val, "val"), (b, "b")):
            assert np.array_equal(got, g[f"{name}_{key}"]), (name, key)


@pytest.mark.parametrize("kind", KINDS)
def test_gradient_lattice_golden(kind):
    g = load("gradient_lattice")
    case = cases.GRADIENT
    dl = 15.0
    p = lambda x, z: 0.5 * x * x / (dl * dl) - z ** 3 / 3.0 / (dl * dl)
    c = cases.lattice_case(kind, case, 20, 20, p=p, dt=case["dt"])
    for st in ("search", "density", "pressure_gradient"):
        c.stage(st)
    P = c.particles()
    for k in ("x", "z", "u", "w"):
        assert np.array_equal(P[k], g[k]), k


@pytest.mark.parametrize("kind", KINDS)
def test_dambreak_1k_golden(kind):
    """Step 1: neighbour lists, A, b and n bit-exact; p within the CG
    tolerance.  Step 10: positions within 1e-9 of the span (the CG solutions
    of the two paths differ within tolerance, and that difference propagates)."""
    g = load("dambreak_1k")
    cfg = scenes.config("dambreak_1k")
    c = cases.make(kind, cfg["env"], cfg["eps"])
    c.add_particles(cfg["particles"])
    c.forward_time(cfg["dt"])
    cnt, ids = c.neighbors()
    assert np.array_equal(cnt, g["step1_count"]) and np.array_equal(ids, g["step1_ids"])
    rp, col, val, b, _ = c.ppe()
    for got, key in ((rp, "rowptr"), (col, "col"), (val, "val"), (b, "b")):
        assert np.array_equal(got, g[f"step1_{key}"]), key
    P = c.particles()
    assert np.array_equal(P["n"], g["step1_n"])
    if kind == "oracle":
        assert np.array_equal(P["p"], g["step1_p"])
    else:
        assert np.abs(P["p"] - g["step1_p"]).max() <= 1e-6 * np.abs(g["step1_p"]).max()
    for _ in range(9):
        c.forward_time(cfg["dt"])
    P = c.particles()
    assert np.array_equal(P["type"], g["step10_type"])
    for k in ("x", "z", "u", "w"):
        ref = g[f"step10_{k}"]
        if kind == "oracle":
            assert np.array_equal(P[k], ref), k
        else:
            assert np.abs(P[k] - ref).max() <= 1e-9 * max(np.abs(ref).max(), 1e-12), k
