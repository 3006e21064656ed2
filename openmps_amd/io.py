"""Particle CSV and scene XML in the reference's formats.

* ``write_csv``: OutputToCsv (Main.cpp:31-67): header ``Type, x, z, u, w, p, n``
  and ``%g``-style values with 6 significant digits (the default precision of
  the C++ stream), or full precision with ``digits=17`` for parity work.
* ``read_csv``: InputFromCsv (Main.cpp:70-183): blank leading lines skipped,
  spaces and tabs removed, columns found by header name.
* ``read_xml``: the ``openmps`` document of the Benchmark generators
  (``condition``, ``environment`` attributes ``value``, ``particles`` CSV),
  Main.cpp:185-274 (LoadParticles, LoadEnvironment, LoadCondition).
* ``write_xml``: the same document layout, as
  Benchmark/DamBreak/generate_koshizukaoka1996.py writes it.
"""
from __future__ import annotations

import io
import xml.etree.ElementTree as ET

import numpy as np

HEADER = "Type, x, z, u, w, p, n"
COLUMNS = ("Type", "x", "z", "u", "w", "p", "n")


def write_csv(particles: dict, path_or_file, digits: int = 6) -> int:
    """Writes particles; returns the number of non-disabled particles."""
    t = np.asarray(particles["type"]).astype(int)
    cols = [np.asarray(particles[k], dtype=np.float64) for k in ("x", "z", "u", "w", "p", "n")]
    fmt = f"%.{digits}g"
    lines = [HEADER]
    for i in range(len(t)):
        lines.append(", ".join([str(t[i])] + [fmt % c[i] for c in cols]))
    text = "\n".join(lines) + "\n"
    if hasattr(path_or_file, "write"):
        path_or_file.write(text)
    else:
        with open(path_or_file, "w") as f:
            f.write(text)
    return int((t != 3).sum())


def read_csv(text: str) -> dict:
    rows = [ln for ln in text.split("\n")]
    k = 0
    while k < len(rows) and rows[k].strip() == "":
        k += 1
    clean = lambda s: s.replace(" ", "").replace("\t", "")
    header = clean(rows[k]).split(",")
    idx = {name: (header.index(name) if name in header else None) for name in COLUMNS}
    data = {name: [] for name in COLUMNS}
    for ln in rows[k + 1:]:
        ln = clean(ln)
        if not ln:
            continue
        items = ln.split(",")
        for name in COLUMNS:
            j = idx[name]
            data[name].append(items[j] if j is not None else "0")
    out = {k2: np.array([float(v) for v in data[k1]], dtype=np.float64)
           for k1, k2 in (("x", "x"), ("z", "z"), ("u", "u"), ("w", "w"), ("p", "p"), ("n", "n"))}
    out["type"] = np.array([int(float(v)) for v in data["Type"]], dtype=np.uint8)
    return out


def read_xml(path: str) -> dict:
    """Condition, environment and particles of an OpenMps input document.

    Keys as LoadCondition (Main.cpp:258-274: startTime, endTime,
    outputInterval, eps), LoadEnvironment (Main.cpp:202-256: l_0,
    minStepCountPerOutput, courant, g, rho, nu, r_eByl_0, minX, minZ, maxX,
    maxZ; the maximum step is outputInterval / minStepCountPerOutput) and
    LoadParticles (Main.cpp:185-199: ``<particles type="csv">``, any other
    type raises as the reference does).  Values may carry blanks (" 0.584").
    Extra environment keys the 2D SPP method does not read (surfaceRatio, c,
    minY, maxY) are kept in ``environment`` but not used."""
    root = ET.parse(path).getroot()
    if root.tag != "openmps":
        raise ValueError(f"not an OpenMps document: root element <{root.tag}>")

    def values(tag):
        node = root.find(tag)
        if node is None:
            raise KeyError(f"openmps.{tag} missing")
        return {child.tag: float(child.get("value").strip()) for child in node
                if isinstance(child.tag, str) and child.get("value") is not None}

    cond = values("condition")
    env = values("environment")
    for k in ("startTime", "endTime", "outputInterval", "eps"):
        if k not in cond:
            raise KeyError(f"openmps.condition.{k} missing")
    node = root.find("particles")
    if node is None or node.get("type") != "csv":
        raise NotImplementedError("Not Implemented!")  # Main.cpp:196-198
    parts = read_csv(node.text or "")
    steps_per_output = int(env["minStepCountPerOutput"])  # std::size_t in Main.cpp:205
    env_args = dict(max_dt=cond["outputInterval"] / steps_per_output,
                    courant=env["courant"], g=env["g"], rho=env["rho"], nu=env["nu"],
                    r_e_by_l0=env["r_eByl_0"], l0=env["l_0"], min_x=env["minX"], min_z=env["minZ"],
                    max_x=env["maxX"], max_z=env["maxZ"])
    return {"condition": cond, "environment": env, "env": env_args, "particles": parts}


def write_xml(path: str, condition: dict, environment: dict, particles: dict, digits: int = 17) -> None:
    """An OpenMps input document: <openmps> with <condition> and
    <environment> children carrying ``value`` attributes and the particles as
    a CSV body of <particles type="csv"> (the layout the Benchmark generators
    write and Main.cpp reads)."""
    root = ET.Element("openmps")
    for tag, d in (("condition", condition), ("environment", environment)):
        node = ET.SubElement(root, tag)
        for k, v in d.items():
            ET.SubElement(node, k).set("value", repr(v) if isinstance(v, float) else str(v))
    buf = io.StringIO()
    write_csv(particles, buf, digits=digits)
    p = ET.SubElement(root, "particles")
    p.set("type", "csv")
    p.text = "\n" + buf.getvalue()
    ET.ElementTree(root).write(path, encoding="utf-8", xml_declaration=True)
