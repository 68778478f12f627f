"""Writes tests/golden/sfmt_radial.json: the reference's published soft-FMT
radial profiles papers/fuzzy-fmt/figs/new-data/radial-wca-{kT}-{n}.dat
(r/sigma and n sigma^3 along +z from a WCA test particle, written by
papers/fuzzy-fmt/figs/new-radial-wca.cpp: 16^3 cell at dx = 0.02, sigma =
epsilon = 1, Xi from papers/fuzzy-fmt/figs/findxi.h).  The three converged
two-column runs (far field back at the bulk density) are kept.  Data only:
the files are read as text."""

import json
import os

REF = "/root/reference/papers/fuzzy-fmt/figs/new-data"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sfmt_radial.json")
cases = []
for kT, n in (("1.0000", "0.60"), ("2.5000", "1.00"), ("10.0000", "1.00")):
    path = os.path.join(REF, f"radial-wca-{kT}-{n}.dat")
    rows = [list(map(float, l.split())) for l in open(path) if l.strip()]
    cases.append({"kT": float(kT), "n_reduced": float(n), "file": os.path.basename(path),
                  "r_over_sigma": [r[0] for r in rows], "n_sigma3": [r[1] for r in rows]})
json.dump({"source": "papers/fuzzy-fmt/figs/new-data (new-radial-wca.cpp output, 800^3)", "cases": cases},
          open(OUT, "w"), indent=0)
print(OUT, len(cases))
