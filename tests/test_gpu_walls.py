"""Moving walls (GPU): the device-side kinematics of mps_set_wall_motion
against the reference's mechanism, host positionWall callbacks
(Computer.hpp:993, :1013-1019) computing the same motion."""
import math
import sys

import numpy as np
import pytest

import cases
from openmps_amd import Computer, make_environment, scenes
from openmps_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def piston_scene():
    """dambreak_1k with the right tank wall (and its dummies) as group 1."""
    cfg = scenes.config("dambreak_1k")
    P = cfg["particles"]
    group = np.where((P["type"] != 0) & (P["x"] > 0.995), 1, 0).astype(np.uint8)
    assert 100 < group.sum() < len(group)
    return cfg, group


def run(cfg, steps, wall_fn=None, motion=None, group=None):
    c = Computer(cfg["eps"], make_environment(cfg["env"]), positionWall=wall_fn)
    c.add_particles(cfg["particles"])
    if motion is not None:
        c.set_wall_motion([motion], group)
    for _ in range(steps):
        c.forward_time(cfg["dt"])
    return c.particles()


def test_translation_matches_callbacks_bitwise():
    cfg, group = piston_scene()
    P0 = cfg["particles"]
    vx = 0.05
    say(f"  {int(group.sum())} moving wall particles, 5 steps")

    def pw(i, t, dt):  # positionWall(i, t, dt): the same arithmetic as the device
        if group[i]:
            return P0["x"][i] + (vx * t + 0.0 * math.sin(0.0 * t + 0.0)), P0["z"][i] + (0.0 * t + 0.0)
        return P0["x"][i], P0["z"][i]

    a = run(cfg, 5, wall_fn=pw)
    b = run(cfg, 5, motion={"vx": vx}, group=group)
    for k in ("x", "z", "u", "w", "p", "n"):
        assert np.array_equal(a[k], b[k]), k
    moved = group == 1
    t = 5 * cfg["dt"]
    assert np.allclose(b["x"][moved] - P0["x"][moved], vx * t, rtol=0, atol=1e-15)
    assert np.allclose(b["u"][moved], vx, rtol=1e-9)


def test_oscillation_positions_and_velocities():
    cfg, group = piston_scene()
    P0 = cfg["particles"]
    amp, omega = 0.002, 2 * math.pi * 5.0
    b = run(cfg, 4, motion={"ax": amp, "omega": omega}, group=group)
    moved = group == 1
    dt = cfg["dt"]
    t = 4 * dt
    want = P0["x"][moved] + amp * math.sin(omega * t)
    assert np.allclose(b["x"][moved], want, rtol=0, atol=1e-15)
    vel = amp * (math.sin(omega * t) - math.sin(omega * (t - dt))) / dt
    assert np.allclose(b["u"][moved], vel, rtol=1e-9, atol=1e-10)
    still = (group == 0) & (P0["type"] != 0)
    assert np.array_equal(b["x"][still], P0["x"][still])


def test_bad_groups_rejected():
    cfg, group = piston_scene()
    c = Computer(cfg["eps"], make_environment(cfg["env"]))
    c.add_particles(cfg["particles"])
    bad = group.copy()
    bad[0] = 2  # no second motion
    with pytest.raises(Exception):
        c.set_wall_motion([{"vx": 1.0}], bad)
    c.set_wall_motion([], group)  # off
    c.forward_time(cfg["dt"])
