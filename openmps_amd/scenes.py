"""Synthetic input scenes in the layout of the reference's scene generators.

Tank walls follow ``Benchmark/DamBreak/generate_koshizukaoka1996.py:59-159``
and ``Benchmark/StaticPressure/generate.py:56-201``: one layer of Wall
particles lining the floor and both side walls, three layers of Dummy
particles behind it, wall tops, and the Dummy corner blocks.  Fluid
particles come first, column by column (x outer, z inner), as in the
generators.  Positions may be jittered uniformly by +/- jitter*l0 from a
seeded numpy PCG64 generator; walls are never jittered.

Particle counts use round(extent / l0) so that an extent that is an exact
multiple of l0 in decimal (0.25 / 0.01) gives that multiple.
"""
from __future__ import annotations

import numpy as np

FLUID, WALL, DUMMY = 0, 1, 2


def _pack(xs, zs, ts) -> dict:
    x = np.asarray(xs, dtype=np.float64)
    z = np.asarray(zs, dtype=np.float64)
    n = len(x)
    return {"x": x, "z": z, "u": np.zeros(n), "w": np.zeros(n), "p": np.zeros(n),
            "n": np.zeros(n), "type": np.asarray(ts, dtype=np.uint8)}


def tank_walls(nw: int, nh: int, l0: float):
    """Wall and Dummy particles of a tank nw * l0 wide with walls nh * l0 high
    (the lattice index layout of generate_koshizukaoka1996.py:76-140)."""
    xs, zs, ts = [], [], []

    def add(i, j, t):
        xs.append(i * l0)
        zs.append(j * l0)
        ts.append(t)

    for i in range(-1, nw + 1):  # floor: wall at z = -1, dummies below
        add(i, -1, WALL)
        for j in (-2, -3, -4):
            add(i, j, DUMMY)
    for j in range(0, nh - 1):  # left wall
        add(-1, j, WALL)
        for i in (-2, -3, -4):
            add(i, j, DUMMY)
    for j in range(0, nh - 1):  # right wall
        add(nw, j, WALL)
        for i in (nw + 1, nw + 2, nw + 3):
            add(i, j, DUMMY)
    for i in range(0, 4):  # wall tops
        add(-(i + 1), nh - 1, WALL)
        add(nw + i, nh - 1, WALL)
    for i in range(1, 4):  # corners
        for j in range(-4, 0):
            add(-1 - i, j, DUMMY)
            add(nw + i, j, DUMMY)
    return xs, zs, ts


def dam_break(tank_w: float, tank_h: float, col_w: float, col_h: float, l0: float,
              jitter: float = 0.0, seed: int = 0) -> tuple[dict, dict]:
    """A water column col_w x col_h at the left of a tank_w x tank_h tank.

    Returns (particles, environment arguments)."""
    nfx, nfz = int(round(col_w / l0)), int(round(col_h / l0))
    nw, nh = int(round(tank_w / l0)), int(round(tank_h / l0))
    i, j = np.meshgrid(np.arange(nfx), np.arange(nfz), indexing="ij")
    fx = (i.ravel() * l0).astype(np.float64)
    fz = (j.ravel() * l0).astype(np.float64)
    if jitter > 0:
        rng = np.random.Generator(np.random.PCG64(seed))
        fx = fx + rng.uniform(-jitter * l0, jitter * l0, fx.shape)
        fz = fz + rng.uniform(-jitter * l0, jitter * l0, fz.shape)
    wx, wz, wt = tank_walls(nw, nh, l0)
    parts = _pack(np.concatenate([fx, wx]), np.concatenate([fz, wz]),
                  np.concatenate([np.zeros(len(fx), dtype=np.uint8), np.asarray(wt, np.uint8)]))
    env = {"min_x": -4 * l0, "min_z": -4 * l0, "max_x": tank_w, "max_z": 2 * tank_h}
    return parts, env


def still_water(width: float, depth: float, l0: float, wall_h: float | None = None,
                jitter: float = 0.0, seed: int = 0) -> tuple[dict, dict]:
    """A tank width wide filled to depth (the static-pressure case)."""
    wall_h = depth * 1.1 if wall_h is None else wall_h
    return dam_break(width, wall_h, width, depth, l0, jitter, seed)


def static_pressure(width: int = 50, height: int = 100, l0: float = 1e-3) -> tuple[dict, dict]:
    """The reference's static-pressure case (Benchmark/StaticPressure/generate.py:
    56-201): width x height fluid particles at rest, a Wall layer with three
    Dummy layers below the floor and beside both walls up to row `height`, and
    the Dummy corner blocks.  Unlike the dam-break tanks the wall tops are
    ordinary wall rows (Wall inside, Dummy outside)."""
    xs, zs, ts = [], [], []

    def add(i, j, t):
        xs.append(i * l0)
        zs.append(j * l0)
        ts.append(t)

    for i in range(width):
        for j in range(height):
            add(i, j, FLUID)
    for i in range(-1, width + 1):  # floor
        add(i, -1, WALL)
        for j in (-2, -3, -4):
            add(i, j, DUMMY)
    for j in range(height + 1):  # left wall
        add(-1, j, WALL)
        for i in (-2, -3, -4):
            add(i, j, DUMMY)
    for j in range(height + 1):  # right wall
        add(width, j, WALL)
        for i in (width + 1, width + 2, width + 3):
            add(i, j, DUMMY)
    for i in range(1, 4):  # corners
        for j in range(-4, 0):
            add(-1 - i, j, DUMMY)
            add(width + i, j, DUMMY)
    parts = _pack(xs, zs, ts)
    # domain: the particles' bounding box (generate.py:190-195)
    env = {"min_x": -4 * l0, "min_z": -4 * l0, "max_x": (width + 3) * l0, "max_z": height * l0}
    return parts, env


def zhou1999(l0: float = 0.02) -> tuple[dict, dict]:
    """Dam break onto a dry deck after Zhou et al. (1999), in the layout of
    Benchmark/DamBreak/generate_zhouetal1999.py:32-151: a tank 1.2 + 2.02 m
    wide with 1 m walls, a 1.2 m wide and 0.6 m high column against the right
    wall (fluid particles first, x = (W - 1 - i) l0 outer, z inner), the same
    wall, dummy, wall-top and corner layers as the other tanks.  Particle
    counts use math.ceil of the float quotients, as the generator does
    (3.22 / 0.02 gives 162 columns).  Domain (-4 l0, -4 l0) to ((W + 4) l0,
    1.2 H l0)."""
    import math
    H = math.ceil(1000e-3 / l0)
    W = math.ceil((1200e-3 + 2020e-3) / l0)
    h = math.ceil(0.6 / l0)
    w = math.ceil(1200e-3 / l0)
    i, j = np.meshgrid(np.arange(w), np.arange(h), indexing="ij")
    fx = ((W - i.ravel() - 1) * l0).astype(np.float64)
    fz = (j.ravel() * l0).astype(np.float64)
    wx, wz, wt = tank_walls(W, H, l0)
    parts = _pack(np.concatenate([fx, wx]), np.concatenate([fz, wz]),
                  np.concatenate([np.zeros(len(fx), dtype=np.uint8), np.asarray(wt, np.uint8)]))
    env = {"min_x": -4 * l0, "min_z": -4 * l0, "max_x": (W + 4) * l0, "max_z": 1.2 * H * l0}
    return parts, env


def lattice(nx: int, nz: int, l0: float, u=None, w=None, p=None) -> dict:
    """nx * nz fluid particles on a square lattice, index i + nx * j (the
    layout of the reference unit tests, e.g. test_ComputerNumberDensity.cpp:82-99)."""
    j, i = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    x = (i.ravel() * l0).astype(np.float64)
    z = (j.ravel() * l0).astype(np.float64)
    s = _pack(x, z, np.zeros(len(x), dtype=np.uint8))
    if u is not None:
        s["u"] = np.asarray(u(x, z), dtype=np.float64) * np.ones_like(x)
    if w is not None:
        s["w"] = np.asarray(w(x, z), dtype=np.float64) * np.ones_like(x)
    if p is not None:
        s["p"] = np.asarray(p(x, z), dtype=np.float64) * np.ones_like(x)
    return s


def batch(count: int = 64, seed: int = 5) -> list[dict]:
    """Uneven independent scenes (a batch of simulations), dam breaks and
    static-pressure columns alternately: fluid particle counts drawn from
    [10k, 200k], l0 from [0.5, 2] mm and the column's height / width from
    [0.3, 2.0] (numpy PCG64, `seed`).  A dam break has its column in a tank 4
    column widths wide with walls 1.2 column heights high, jittered +/-0.02
    l0; a static-pressure scene is the reference's layout (static_pressure:
    the column at rest between walls as high as the column).  r_e = 2.1 l0,
    r_lap = 3.1 l0, steps up to 1e-4 s (CFL rule).  Returns config
    dictionaries like config(), with "kind" set."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k in range(count):
        nfluid = int(rng.integers(10_000, 200_001))
        l0 = float(rng.uniform(0.5e-3, 2e-3))
        aspect = float(rng.uniform(0.3, 2.0))
        nx = max(4, int(round(np.sqrt(nfluid / aspect))))
        nz = max(4, int(round(nfluid / nx)))
        col_w, col_h = nx * l0, nz * l0
        if k % 2 == 0:
            kind = "dambreak"
            parts, dom = dam_break(4 * col_w, 1.2 * col_h, col_w, col_h, l0, 0.02, seed * 1000 + k)
        else:
            kind = "static"
            parts, dom = static_pressure(nx, nz, l0)
        env = dict(max_dt=1e-4, courant=COURANT, g=G, rho=RHO, nu=NU, r_e_by_l0=2.1, l0=l0,
                   r_lap_by_l0=3.1, **dom)
        out.append({"name": f"batch{seed}_{k}", "particles": parts, "env": env, "dt": None, "eps": 1e-8,
                    "l0": l0, "aspect": aspect, "kind": kind, "depth": col_h})
    return out


# Named workloads: the configurations the benchmarks and parity tests use.
RHO, NU, G, COURANT = 998.2, 1e-6, 9.8, 0.1


def config(name: str, seed: int = 1) -> dict:
    """Scene + environment + run parameters of a named workload.

    dambreak_1k: 1.0 x 0.6 m tank, 0.25 x 0.5 m column, l0 = 10 mm, no jitter,
                 r_e = 2.1 l0, r_lap = 3.1 l0, dt = 1e-3, eps = 1e-10.
    dambreak_125k: same geometry at l0 = 1 mm, jitter +/-0.02 l0, dt = 1e-4.
    dambreak_1m: 2.0 x 1.2 m tank, 0.5 x 1.0 m column, l0 = 0.7 mm, jitter
                 +/-0.02 l0, dt = 1e-4.
    still_1m:    0.5 m wide tank filled 1.0 m deep at l0 = 0.7 mm, dt = 2e-5.
    static_pressure: the reference's static-pressure case, 50 x 100 fluid
                 particles at l0 = 1 mm, r_e = 2.4 l0, steps up to 5e-6 s.
    zhou1999:    the Zhou et al. (1999) dam break of the reference's generator,
                 l0 = 20 mm, r_e = 2.4 l0, CFL steps up to 5e-4 s.
    Each accepts a suffix ``_ref`` for the reference's single radius
    (r_lap = r_e = 2.1 l0).
    """
    single = name.endswith("_ref")
    base = name[:-4] if single else name
    if base == "dambreak_1k":
        l0, dt, eps, jit = 0.01, 1e-3, 1e-10, 0.0
        parts, dom = dam_break(1.0, 0.6, 0.25, 0.5, l0)
    elif base == "dambreak_125k":
        l0, dt, eps, jit = 0.001, 1e-4, 1e-10, 0.02
        parts, dom = dam_break(1.0, 0.6, 0.25, 0.5, l0, jit, seed)
    elif base == "dambreak_1m":
        l0, dt, eps, jit = 0.0007, 1e-4, 1e-8, 0.02
        parts, dom = dam_break(2.0, 1.2, 0.5, 1.0, l0, jit, seed)
    elif base == "still_1m":
        # the column starts at rest with p = 0 (as the reference's static-pressure
        # generator, Benchmark/StaticPressure/generate.py, which steps at 5e-6 s);
        # 1e-4 s lets the 1 m column's start-up compression grow unstable
        l0, dt, eps, jit = 0.0007, 2e-5, 1e-8, 0.0
        parts, dom = still_water(0.5, 1.0, l0)
    elif base == "static_pressure":
        # Benchmark/StaticPressure/generate.py:8-31: l0 = 1 mm, r_e = 2.4 l0 (one
        # radius), nu = 1.004e-6, maximum step outputInterval / minStepCountPerOutput
        # = 0.0005 / 100 (Main.cpp:234), eps = 1e-10
        parts, dom = static_pressure()
        env = dict(max_dt=5e-6, courant=COURANT, g=G, rho=RHO, nu=1.004e-6, r_e_by_l0=2.4,
                   l0=1e-3, r_lap_by_l0=2.4, **dom)
        return {"name": name, "particles": parts, "env": env, "dt": 5e-6, "eps": 1e-10}
    elif base == "zhou1999":
        # Benchmark/DamBreak/generate_zhouetal1999.py:11-28: l0 = 20 mm, r_e = 2.4 l0
        # (one radius), nu = 1.004e-6, maximum step outputInterval /
        # minStepCountPerOutput = 0.005 / 10, eps = 1e-10
        parts, dom = zhou1999(0.02)
        env = dict(max_dt=5e-4, courant=COURANT, g=G, rho=RHO, nu=1.004e-6, r_e_by_l0=2.4,
                   l0=0.02, r_lap_by_l0=2.4, **dom)
        return {"name": name, "particles": parts, "env": env, "dt": None, "eps": 1e-10}
    elif base.startswith("dambreak_l0_"):
        l0 = float(base[len("dambreak_l0_"):])
        dt, eps, jit = 1e-4, 1e-10, 0.02
        parts, dom = dam_break(1.0, 0.6, 0.25, 0.5, l0, jit, seed)
    else:
        raise KeyError(name)
    env = dict(max_dt=dt, courant=COURANT, g=G, rho=RHO, nu=NU, r_e_by_l0=2.1, l0=l0,
               r_lap_by_l0=(2.1 if single else 3.1), **dom)
    return {"name": name, "particles": parts, "env": env, "dt": dt, "eps": eps}
