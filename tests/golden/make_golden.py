"""Extract the reference-held golden values used by the oracle tests.

Reads the conservation traces the reference ships as data
(paper/data/conservation-*.csv, written by FluidSim2D::saveStats,
src/FluidSim2D.cpp:738-750) and stores their first rows in
tests/golden/paper_conservation.json.  Run in a container that has the
reference checkout mounted at /root/reference; the tests read only the JSON.
"""
import json
import os
import sys

REF = os.environ.get("FS2D_REFERENCE", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "paper_conservation.json")
NROWS = 8


def rows(path):
    out = []
    with open(path) as f:
        next(f)  # "Total Volume, Total Energy, Total Energy (Particle)"
        for line in f:
            vals = [float(x) for x in line.split(",") if x.strip()]
            out.append(vals)
            if len(out) == NROWS:
                break
    return out


def main():
    data = {
        "source": "paper/data/conservation-*.csv (columns: Total Volume, Total Energy, "
                  "Total Energy (Particle)); one row per frame, printed with %f",
        "scene": {"n": 128, "dx": 0.02, "dt": 0.033, "rho": 997.0, "gravity": [0.0, -9.81],
                  "particles_per_cell_sqrt": 2, "cells": "demo/App.cpp:145-158 triangle i+j<96"},
        "files": {},
    }
    names = ["conservation-semilag-state1.csv"] + [
        "conservation-pic-alpha%d-state1.csv" % a for a in (50, 60, 70, 80, 90, 100)]
    for n in names:
        data["files"][n] = rows(os.path.join(REF, "paper", "data", n))
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
