"""Writes tests/golden/sw_radial.json: the reference's published square-well
liquid profiles papers/square-well-fluid/data/radial-sw-{kT}-1.30-{ff}.dat
(columns r/sigma and the packing fraction n pi sigma^3/6 along +z from the
test particle; written by papers/square-well-fluid/figs/new-radial-sw.cpp:
18^3 cell, dx = 0.1, R = 1, epsilon = 1, lambda = 1.3).  Data only: the
reference's files are read as text, nothing of it is run."""

import json
import os

REF = "/root/reference/papers/square-well-fluid/data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sw_radial.json")

cases = []
for kT in ("3.00", "5.00", "10.00"):
    for ff in ("0.20", "0.30"):
        path = os.path.join(REF, f"radial-sw-{kT}-1.30-{ff}.dat")
        rows = [list(map(float, l.split()))[:2] for l in open(path) if l.strip()]
        cases.append({"kT": float(kT), "lambda": 1.3, "ff": float(ff), "file": os.path.basename(path),
                      "r_over_sigma": [r[0] for r in rows], "eta": [r[1] for r in rows]})
json.dump({"source": "papers/square-well-fluid/data (new-radial-sw.cpp output)", "cases": cases},
          open(OUT, "w"), indent=0)
print(OUT, len(cases))
