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


# An input document in the reference's layout (Benchmark/Sample/Sample.xml,
# Benchmark/DamBreak/generate_koshizukaoka1996.py): comments, blanks inside
# values, keys the 2D SPP method does not read, stray text, a CSV body.
SAMPLE_XML = """<?xml version="1.0" encoding="utf-8"?>
<openmps>
  <condition>
    <startTime value="0" /> <!-- start -->
    <endTime value="0.03" />
    <outputInterval value="0.01" />
    <eps value="1e-10" />
  </condition>
  <environment>
    <l_0 value="0.02" />
    <minStepCountPerOutput value="10" />
    <courant value="0.1" />
    <g value="9.8" />
    <rho value="998.20" />
    <nu value="1.004e-6" />
    <r_eByl_0 value="2.4" />
    <surfaceRatio value="0.97" />
    <minX value="-0.08" />
    <minY value="-0.000" />
    <minZ value="-0.08" />
    <maxX value=" 0.584" />
    <maxY value=" 0.000" />
    <maxZ value=" 0.6" />
    <c value="1.5" /> -->
  </environment>
  <particles type="csv">
Type, x, z, u, w, p, n
0, 0.1, 0.2, 0.5, -1.5, 3, 4
2,-0.04,  -0.06, 0,0,0,0
1, 0.25, 0, 0, 0, 0, 0
  </particles>
</openmps>
"""


def test_read_xml_reference_layout(tmp_path):
    """LoadCondition / LoadEnvironment / LoadParticles (Main.cpp:185-274) on a
    document in the reference's layout: maxDt = outputInterval /
    minStepCountPerOutput, blanks inside values, unused keys ignored."""
    p = tmp_path / "in.xml"
    p.write_text(SAMPLE_XML)
    doc = mio.read_xml(str(p))
    assert doc["condition"] == {"startTime": 0.0, "endTime": 0.03, "outputInterval": 0.01, "eps": 1e-10}
    e = doc["env"]
    assert e["max_dt"] == 0.01 / 10
    assert (e["l0"], e["r_e_by_l0"], e["courant"], e["g"], e["rho"], e["nu"]) == (0.02, 2.4, 0.1, 9.8, 998.2, 1.004e-6)
    assert (e["min_x"], e["min_z"], e["max_x"], e["max_z"]) == (-0.08, -0.08, 0.584, 0.6)
    s = doc["particles"]
    assert s["type"].tolist() == [0, 2, 1]
    assert s["x"].tolist() == [0.1, -0.04, 0.25] and s["z"].tolist() == [0.2, -0.06, 0.0]
    assert s["u"].tolist() == [0.5, 0.0, 0.0] and s["w"].tolist() == [-1.5, 0.0, 0.0]
    assert s["p"].tolist() == [3.0, 0.0, 0.0] and s["n"].tolist() == [4.0, 0.0, 0.0]


def test_read_xml_rejects_other_particle_formats(tmp_path):
    p = tmp_path / "in.xml"
    p.write_text(SAMPLE_XML.replace('type="csv"', 'type="binary"'))
    with pytest.raises(NotImplementedError):
        mio.read_xml(str(p))


def test_write_xml_round_trip(tmp_path):
    cfg = scenes.config("dambreak_1k")
    e = cfg["env"]
    cond = {"startTime": 0.0, "endTime": 0.01, "outputInterval": 0.005, "eps": 1e-10}
    env = {"l_0": e["l0"], "minStepCountPerOutput": 5, "courant": e["courant"], "g": e["g"], "rho": e["rho"],
           "nu": e["nu"], "r_eByl_0": e["r_e_by_l0"], "minX": e["min_x"], "minZ": e["min_z"],
           "maxX": e["max_x"], "maxZ": e["max_z"]}
    p = tmp_path / "scene.xml"
    mio.write_xml(str(p), cond, env, cfg["particles"])
    doc = mio.read_xml(str(p))
    assert doc["condition"] == cond
    assert doc["env"]["max_dt"] == 0.005 / 5
    for k in ("x", "z", "u", "w", "p", "n", "type"):
        assert np.array_equal(doc["particles"][k], cfg["particles"][k]), k


@pytest.mark.gpu
def test_xml_driven_run_matches_direct_stepping(tmp_path):
    """openmps_amd.run (the Main.cpp:288-400 loop: ForwardTime() until the next
    output time, particles_%05d.csv per interval) on a coarse dam break equals
    the same CFL steps taken directly through the Computer."""
    from openmps_amd import CreateComputer, make_environment
    from openmps_amd import run as mrun
    s, dom = scenes.dam_break(1.0, 0.6, 0.25, 0.5, 0.02)
    cond = {"startTime": 0.0, "endTime": 0.03, "outputInterval": 0.01, "eps": 1e-10}
    env = {"l_0": 0.02, "minStepCountPerOutput": 4, "courant": 0.1, "g": 9.8, "rho": 998.2, "nu": 1e-6,
           "r_eByl_0": 2.1, "minX": dom["min_x"], "minZ": dom["min_z"], "maxX": dom["max_x"], "maxZ": dom["max_z"]}
    p = tmp_path / "scene.xml"
    mio.write_xml(str(p), cond, env, s)
    out = tmp_path / "result"
    lines = []
    res = mrun.run(str(p), str(out), say=lines.append)
    assert res["error"] is None and lines[-1] == "finished"
    assert [os.path.basename(f) for f in res["files"]] == [f"particles_{k:05d}.csv" for k in range(4)]
    assert lines[1].startswith("#00000: t=  0.0000 (00000), ") and lines[4].startswith("#00003: t=")
    doc = mio.read_xml(str(p))
    c = CreateComputer(1e-10, make_environment(doc["env"]))
    c.AddParticles(doc["particles"])
    steps, next_t = 0, 0.0
    for k in range(1, 4):
        next_t += 0.01  # accumulated as Main.cpp:366 does
        while c.time()[0] < next_t:
            c.ForwardTime()
            steps += 1
        buf = io.StringIO()
        mio.write_csv(c.Particles(), buf)
        assert (out / f"particles_{k:05d}.csv").read_text() == buf.getvalue(), k
    assert steps == res["iterations"] >= 12  # maxDt = 0.01 / 4
