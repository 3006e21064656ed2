"""Particle CSV and scene XML in the reference's formats.

* ``write_csv``: OutputToCsv (Main.cpp:31-67): header ``Type, x, z, u, w, p, n``
  and ``%g``-style values with 6 significant digits (the default precision of
  the C++ stream), or full precision with ``digits=17`` for parity work.
* ``read_csv``: InputFromCsv (Main.cpp:70-183): blank leading lines skipped,
  spaces and tabs removed, columns found by header name.
* ``read_xml``: the ``openmps`` document of the Benchmark generators
  (``condition``, ``environment`` attributes ``value``, ``particles`` CSV),
  Main.cpp:202-274.
"""
from __future__ import annotations

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
    """Condition, environment and particles of an OpenMps input document."""
    root = ET.parse(path).getroot()

    def values(tag):
        node = root.find(tag)
        return {child.tag: float(child.get("value")) for child in node} if node is not None else {}

    cond = values("condition")
    env = values("environment")
    parts = read_csv(root.find("particles").text)
    # LoadEnvironment, Main.cpp:234: maxDt = outputInterval / minStepCountPerOutput
    env_args = dict(max_dt=cond["outputInterval"] / env["minStepCountPerOutput"],
                    courant=env["courant"], g=env["g"], rho=env["rho"], nu=env["nu"],
                    r_e_by_l0=env["r_eByl_0"], l0=env["l_0"], min_x=env["minX"], min_z=env["minZ"],
                    max_x=env["maxX"], max_z=env["maxZ"])
    return {"condition": cond, "env": env_args, "particles": parts}
