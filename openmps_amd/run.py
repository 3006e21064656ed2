"""XML-driven simulation run: the loop of the reference's ``main``
(Main.cpp:288-400) over the MI355X Computer.

The input is an OpenMps document (condition, environment, particles as CSV;
``io.read_xml``).  Walls stay at their initial positions (Main.cpp:305-318;
the library's default static walls).  The initial state is written to
``<out>/particles_<k>.csv`` with k = ceil(startTime / outputInterval), then
for each of ceil((endTime - startTime) / outputInterval) output intervals the
computer advances with the CFL step (``ForwardTime()``) until its clock
reaches the next output time and the state is written to the next file
(``particles_%05d.csv``, Main.cpp:34), with the reference's progress line.  A
pressure solve that does not converge prints the reference's error block and
ends the run (Main.cpp:382-390); a cell overflow propagates, as it does there.

Usage:  python -m openmps_amd.run input.xml [--out result] [--solver PRESET]
Without --solver the pressure solve is the reference's FP64 CG with its own
stop rule and the document's eps.
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

from . import io as mio


def _stamp(t_computer, iteration, output_count, count, elapsed) -> str:
    """Main.cpp:334: "#%3$05d: t=%1$8.4lf (%2$05d), %10$12d particles, @ %4$02d/%5$02d
    %6$02d:%7$02d:%8$02d (%9$8.2lf)"."""
    lt = time.localtime()
    return (f"#{output_count:05d}: t={t_computer:8.4f} ({iteration:05d}), {count:12d} particles, "
            f"@ {lt.tm_mon:02d}/{lt.tm_mday:02d} {lt.tm_hour:02d}:{lt.tm_min:02d}:{lt.tm_sec:02d} "
            f"({elapsed:8.2f})")


def run(path: str, out_dir: str = "result", solver: str | None = None, digits: int = 6,
        say=print) -> dict:
    """Runs the document at `path`; returns {"files": [...], "iterations": k,
    "outputs": n, "error": message or None}."""
    from . import CreateComputer, ComputerException, make_environment, solvers
    doc = mio.read_xml(path)
    cond = doc["condition"]
    say(f"Input XML file: {path}")
    env = make_environment(doc["env"])
    computer = CreateComputer(cond["eps"], env)
    if solver is not None:
        solvers.configure(computer, solver, cond["eps"])
    computer.AddParticles(doc["particles"])
    os.makedirs(out_dir, exist_ok=True)
    t_start = time.perf_counter()
    interval = cond["outputInterval"]
    offset = int(math.ceil(cond["startTime"] / interval))
    files = []

    def output(k):
        name = os.path.join(out_dir, f"particles_{k:05d}.csv")
        count = mio.write_csv(computer.Particles(), name, digits=digits)
        files.append(name)
        return count

    say(_stamp(cond["startTime"], 0, offset, output(offset), time.perf_counter() - t_start))
    next_t = 0.0
    iteration = 0
    end_count = int(math.ceil((cond["endTime"] - cond["startTime"]) / interval))
    error = None
    for k in range(1, end_count + 1):
        t_computer = computer.GetEnvironment().T()
        try:
            next_t += interval
            while t_computer < next_t:
                computer.ForwardTime()
                t_computer = computer.GetEnvironment().T()
                iteration += 1
            count = output(k + offset)
            t_computer += cond["startTime"]
            say(_stamp(t_computer, iteration, k + offset, count, time.perf_counter() - t_start))
        except ComputerException as ex:
            say("!!!!ERROR!!!!")
            say(f"#{k + offset}: t={t_computer!r} ({iteration})")
            say(str(ex))
            error = str(ex)
            break
    say("finished")
    return {"files": files, "iterations": iteration, "outputs": len(files), "error": error}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("input", help="OpenMps XML document")
    ap.add_argument("--out", default="result", help="output folder (Main.cpp writes result/)")
    ap.add_argument("--solver", default=None, help="pressure solver preset (openmps_amd.solvers)")
    ap.add_argument("--digits", type=int, default=6,
                    help="significant digits of the CSV values (6: the C++ stream default)")
    a = ap.parse_args(argv)
    run(a.input, a.out, a.solver, a.digits)
    return 0


if __name__ == "__main__":
    sys.exit(main())
