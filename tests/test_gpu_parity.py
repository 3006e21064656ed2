"""HIP path against the CPU oracle, stage by stage (all tests need a GPU).

Bar: neighbour lists, cell disabling, n, nWithoutSpp, ECS, the explicit
update, the assembled PPE (A, b, x0), the pressure write-back, the pressure
gradient and DS are bit-exact (FP64, integer).  The CG solution agrees within
the solver tolerance; after comparing it, the GPU solution is written into
the oracle so that the stages after the solve are again compared bit for bit.
"""
import sys

import numpy as np
import pytest

import cases
from openmps_amd import scenes

pytestmark = pytest.mark.gpu

STAGES_A = ("search", "density", "error_correction", "explicit", "density", "save_x",
            "ppe_assemble")
STAGES_B = ("pressure_writeback", "pressure_gradient", "dynamic_stabilize")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_particles(g, o, what):
    P, Q = g.particles(), o.particles()
    for k in ("x", "z", "u", "w", "p", "n", "type"):
        assert np.array_equal(P[k], Q[k], equal_nan=(k != "type")), \
            f"{what}: particle field {k} differs"


def same_aux(g, o, names, what):
    for k in names:
        assert np.array_equal(g.aux(k), o.aux(k), equal_nan=True), f"{what}: {k} differs"


def compare_stage(g, o, st):
    if st == "search":
        cg, ig = g.neighbors()
        co, io = o.neighbors()
        assert np.array_equal(cg, co), "neighbour counts differ"
        assert np.array_equal(ig, io), "neighbour ids differ"
        assert np.array_equal(g.particles()["type"], o.particles()["type"])
    elif st == "density":
        same_particles(g, o, st)
        same_aux(g, o, ["n_without_spp"], st)
    elif st == "error_correction":
        same_aux(g, o, ["ecs"], st)
    elif st == "explicit":
        same_particles(g, o, st)
        same_aux(g, o, ["du_x", "du_z"], st)
    elif st == "save_x":
        same_aux(g, o, ["saved_x", "saved_z"], st)
    elif st == "ppe_assemble":
        A, B = g.ppe(), o.ppe()
        for a, b, nm in zip(A, B, ("rowptr", "col", "val", "b", "x0")):
            assert np.array_equal(a, b), f"PPE {nm} differs"
    else:
        same_particles(g, o, st)
        if st != "pressure_writeback":
            same_aux(g, o, ["du_x", "du_z"], st)


def run_parity(name, steps, particles=None, env_over=None, keep_order=False):
    cfg = scenes.config(name)
    env = dict(cfg["env"], **(env_over or {}))
    parts = cfg["particles"] if particles is None else particles
    g = cases.make("gpu", env, cfg["eps"])
    if keep_order:
        g.set_storage_order(keep=True)
    o = cases.make("oracle", env, cfg["eps"])
    g.add_particles(parts)
    o.add_particles(parts)
    dt = cfg["dt"]
    t = 0.0
    for step in range(steps):
        t += dt
        g.set_time(t, dt)
        o.set_time(t, dt)
        for st in STAGES_A:
            g.stage(st)
            o.stage(st)
            compare_stage(g, o, st)
        g.stage("ppe_solve")
        o.stage("ppe_solve")
        xg, xo = g.aux("ppe_x"), o.aux("ppe_x")
        rel = np.linalg.norm(xg - xo) / max(np.linalg.norm(xo), 1e-300)
        sg, so = g.cg_stats(), o.cg_stats()
        say(f"  {name} step {step}: CG iters gpu {sg['iterations']} cpu {so['iterations']}, "
            f"|x_gpu - x_cpu| / |x_cpu| = {rel:.2e}")
        assert rel < 1e-6
        assert sg["rr"] < cfg["eps"] ** 2 * sg["rr0"]
        o.set_aux("ppe_x", xg)
        for st in STAGES_B:
            g.stage(st)
            o.stage(st)
            compare_stage(g, o, st)
    return g, o


def test_device_sqrt_div_correctly_rounded():
    """The kernels' FP64 sqrt and division round as the host's (IEEE)."""
    from openmps_amd import device_math_check
    rng = np.random.Generator(np.random.PCG64(3))
    a = np.concatenate([rng.uniform(0, 1e-3, 400000) ** 2, rng.uniform(0, 4.0, 400000),
                        np.exp(rng.uniform(-700, 700, 200000)),
                        np.array([0.0, 1.0, 2.0, 4.0, 1e-320, 5e-324, np.inf])])
    b = rng.uniform(1e-6, 1.0, len(a))
    s, d = device_math_check(a, b)
    assert np.array_equal(s, np.sqrt(a))
    assert np.array_equal(d, a / b)


@pytest.mark.parametrize("name", ["dambreak_1k", "dambreak_1k_ref"])
@pytest.mark.parametrize("storage", ["cell", "keep"])
def test_stage_parity_dambreak(name, storage):
    """Both storage orders: particles permuted into cell order at every search
    (the default) and kept in the caller's order."""
    say(f"parity {name}, storage order {storage}")
    run_parity(name, steps=6, keep_order=(storage == "keep"))


def test_stage_parity_jittered():
    s, _ = scenes.dam_break(1.0, 0.6, 0.25, 0.5, 0.01, jitter=0.2, seed=11)
    say("parity jittered dambreak_1k")
    run_parity("dambreak_1k", steps=4, particles=s)


def test_stage_parity_shuffled():
    """Particles in a random order: consecutive particles (one wavefront)
    are scattered over the whole scene, so the neighbour walks share no
    cells and the lists are written in index order, not spatial order."""
    s, _ = scenes.dam_break(1.0, 0.6, 0.25, 0.5, 0.01, jitter=0.2, seed=11)
    perm = np.random.Generator(np.random.PCG64(5)).permutation(len(s["x"]))
    s = {k: v[perm] for k, v in s.items()}
    say("parity shuffled jittered dambreak_1k")
    run_parity("dambreak_1k", steps=2, particles=s)


@pytest.mark.parametrize("order", ["lattice", "shuffled"])
def test_search_parity_large_cells(order):
    """r_e = 5 l0 (the reference's HS test radius, NL = 6 l0): a cell holds
    36 lattice sites, more entries than the cell sort keeps in registers, so
    its index order is restored by the insertion sort."""
    case = cases.HS
    s = scenes.lattice(35, 35, case["env"]["l0"])
    if order == "shuffled":
        perm = np.random.Generator(np.random.PCG64(9)).permutation(len(s["x"]))
        s = {k: v[perm] for k, v in s.items()}
    g = cases.make("gpu", case["env"], case["eps"])
    o = cases.make("oracle", case["env"], case["eps"])
    g.add_particles(s)
    o.add_particles(s)
    say(f"search parity, r_e = 5 l0, {order}")
    g.stage("search")
    o.stage("search")
    compare_stage(g, o, "search")


def test_stage_parity_125k():
    """One step at the 125k-particle configuration (l0 = 1 mm, jittered)."""
    say("parity dambreak_125k (1 step)")
    run_parity("dambreak_125k", steps=1)


def test_forward_time_matches_stages():
    """ForwardTime(dt) (fused kernels) equals the stage-by-stage sequence."""
    cfg = scenes.config("dambreak_1k")
    g = cases.make("gpu", cfg["env"], cfg["eps"])
    o = cases.make("oracle", cfg["env"], cfg["eps"])
    g.add_particles(cfg["particles"])
    o.add_particles(cfg["particles"])
    for step in range(3):
        g.forward_time(cfg["dt"])
        o.forward_time(cfg["dt"])
        cg, ig = g.neighbors()
        co, io = o.neighbors()
        assert np.array_equal(ig, io)
        P, Q = g.particles(), o.particles()
        for k in ("x", "z", "u", "w"):
            scale = max(np.abs(Q[k]).max(), 1e-12)
            assert np.abs(P[k] - Q[k]).max() <= 1e-9 * scale, k
        assert np.abs(P["p"] - Q["p"]).max() <= 1e-6 * np.abs(Q["p"]).max()
        assert g.time() == o.time()


def test_cfl_forward_time():
    """ForwardTime() with DetermineDt (Computer.hpp:759-777)."""
    cfg = scenes.config("dambreak_1k")
    g = cases.make("gpu", cfg["env"], cfg["eps"])
    o = cases.make("oracle", cfg["env"], cfg["eps"])
    g.add_particles(cfg["particles"])
    o.add_particles(cfg["particles"])
    assert g.determine_dt() == o.determine_dt()  # u = 0: MaxDt
    for _ in range(3):
        dg = g.forward_time(None)
        do = o.forward_time(None)
        assert dg == pytest.approx(do, rel=1e-9)


def test_out_of_grid_particles_disabled():
    cfg = scenes.config("dambreak_1k")
    s = {k: v.copy() for k, v in cfg["particles"].items()}
    s["x"][:5] = [-1.0, 50.0, np.nan, np.inf, 0.3]
    s["z"][:5] = [0.1, 0.1, 0.1, 0.1, -5.0]
    g, o = run_parity("dambreak_1k", steps=1, particles=s)
    assert np.all(g.particles()["type"][:5] == 3)


def test_empty_scene():
    cfg = scenes.config("dambreak_1k")
    g = cases.make("gpu", cfg["env"], cfg["eps"])
    g.forward_time(cfg["dt"])
    assert g.size() == 0


def test_cell_overflow_raises():
    from openmps_amd import GridException
    import oracle as om
    cfg = scenes.config("dambreak_1k")
    s = scenes._pack(np.full(40, 0.1), np.full(40, 0.1), np.zeros(40, dtype=np.uint8))
    g = cases.make("gpu", cfg["env"], cfg["eps"])
    o = cases.make("oracle", cfg["env"], cfg["eps"])
    g.add_particles(s)
    o.add_particles(s)
    with pytest.raises(GridException, match="Too many particle in a block"):
        g.forward_time(cfg["dt"])
    with pytest.raises(om.CellOverflow):
        o.forward_time(cfg["dt"])


def test_cg_not_converged_raises():
    """Inconsistent singular system: CG runs N iterations and fails
    (Computer.hpp:1424-1428); eps = 0 converges immediately (:1386-1389)."""
    from openmps_amd import ComputerException
    import oracle as om
    A = np.diag([1.0, 0.0])
    rp, col, val = cases.csr_from_dense(A)
    b = np.array([1.0, 1.0])
    for kind, exc in (("gpu", ComputerException), ("oracle", om.CgNotConverged)):
        c = cases.make(kind, cases.CG["env"], 1e-7)
        c.set_ppe(rp, col, val, b, np.zeros(2))
        with pytest.raises(exc):
            c.stage("ppe_solve")
        c0 = cases.make(kind, cases.CG["env"], 0.0)
        c0.set_ppe(rp, col, val, b, np.zeros(2))
        c0.stage("ppe_solve")
        assert c0.cg_stats()["iterations"] == 0


def test_add_particles_twice():
    """Appending in two calls equals one call (AddParticles, Computer.hpp:1754-1777)."""
    cfg = scenes.config("dambreak_1k")
    s = cfg["particles"]
    n = len(s["x"])
    half = {k: v[: n // 2] for k, v in s.items()}
    rest = {k: v[n // 2:] for k, v in s.items()}
    g = cases.make("gpu", cfg["env"], cfg["eps"])
    g.add_particles(half)
    g.add_particles(rest)
    o = cases.make("oracle", cfg["env"], cfg["eps"])
    o.add_particles(s)
    g.forward_time(cfg["dt"])
    o.forward_time(cfg["dt"])
    assert np.array_equal(g.neighbors()[1], o.neighbors()[1])


def test_wall_callbacks():
    """Moving walls through positionWall / positionWallPre (Computer.hpp:993,
    :1013-1019) equal the oracle with the same target positions."""
    from openmps_amd import Computer, make_environment
    cfg = scenes.config("dambreak_1k")
    s = cfg["particles"]
    calls = {"pre": 0}

    def pre(t, dt):
        calls["pre"] += 1

    def wall(i, t, dt):
        return s["x"][i] + 0.001 * t, s["z"][i]

    g = Computer(cfg["eps"], make_environment(cfg["env"]), wall, pre)
    g.add_particles(s)
    o = cases.make("oracle", cfg["env"], cfg["eps"])
    o.add_particles(s)
    dt = cfg["dt"]
    for step in range(2):
        t = (step + 1) * dt
        o.set_wall_positions(s["x"] + 0.001 * t, s["z"])
        g.forward_time(dt)
        o.forward_time(dt)
        assert np.array_equal(g.neighbors()[1], o.neighbors()[1])
        P, Q = g.particles(), o.particles()
        wallmask = P["type"] != 0
        assert np.array_equal(P["x"][wallmask], Q["x"][wallmask])
        assert np.array_equal(P["u"][wallmask], Q["u"][wallmask])
    assert calls["pre"] == 2
