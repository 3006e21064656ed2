"""The C++ facade (include/mps_computer.hpp) and the CSV / XML formats."""
import io
import os
import subprocess

import numpy as np
import pytest

import cases
from openmps_amd import io as mio
from openmps_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "facade_main")


def test_csv_round_trip():
    cfg = scenes.config("dambreak_1k")
    s = cfg["particles"]
    buf = io.StringIO()
    count = mio.write_csv(s, buf, digits=17)
    assert count == len(s["x"])
    text = buf.getvalue()
    assert text.splitlines()[0] == "Type, x, z, u, w, p, n"
    back = mio.read_csv("\n\n" + text)
    for k in ("x", "z", "u", "w", "p", "n", "type"):
        assert np.array_equal(back[k], s[k]), k


def test_csv_default_precision_matches_stream_format():
    s = scenes._pack([0.123456789, 1e-7], [2.0, -3.5], [0, 1])
    buf = io.StringIO()
    mio.write_csv(s, buf)
    lines = buf.getvalue().splitlines()
    assert lines[1] == "0, 0.123457, 2, 0, 0, 0, 0"  # std::ostream default: 6 significant
    assert lines[2] == "1, 1e-07, -3.5, 0, 0, 0, 0"


def test_read_csv_missing_columns_default_to_zero():
    s = mio.read_csv("Type,x,z\n0, 1.5, 2\n2,3,4\n")
    assert s["x"].tolist() == [1.5, 3.0] and s["u"].tolist() == [0.0, 0.0]
    assert s["type"].tolist() == [0, 2]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["callbacks", "static"])
def test_cpp_facade_matches_python_path(tmp_path, mode):
    """Main.cpp-style C++ driver over the facade == the same steps through the
    Python mirror (same library, deterministic kernels => identical bits)."""
    cfg = scenes.config("dambreak_1k")
    s = cfg["particles"]
    inp, out = tmp_path / "in.csv", tmp_path / "out.csv"
    mio.write_csv(s, str(inp), digits=17)
    e = cfg["env"]
    assert e["r_lap_by_l0"] != e["r_e_by_l0"]
    e = dict(e, r_lap_by_l0=e["r_e_by_l0"])  # the facade's Environment has the reference's single radius
    steps = 3
    args = [DRIVER, str(inp), str(out), str(steps), repr(cfg["dt"]), repr(cfg["eps"]),
            *[repr(e[k]) for k in ("max_dt", "courant", "g", "rho", "nu", "r_e_by_l0", "l0",
                                   "min_x", "min_z", "max_x", "max_z")]]
    if mode == "static":
        args.append("static")
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    if mode == "callbacks":
        assert f"positionWallPre calls: {steps}" in r.stdout
    got = mio.read_csv(out.read_text())
    g = cases.make("gpu", e, cfg["eps"])
    g.add_particles(mio.read_csv(inp.read_text()))
    for _ in range(steps):
        g.forward_time(cfg["dt"])
    P = g.particles()
    for k in ("x", "z", "u", "w", "p", "n", "type"):
        assert np.array_equal(got[k], P[k]), k
