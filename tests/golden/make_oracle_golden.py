"""Writes the FP64 oracle's converged minimizations used as golden values by the GPU tests.

    python tests/golden/make_oracle_golden.py [c1] [c2]

c1: hard-sphere fluid (R=1, kT=1, eta=0.3) against a planar hard wall
    (Vext = inf for x < R) in a periodic 6.4^3 cell, 64^3 grid.
c2: the same fluid at eta=0.4 around a spherical hard solute (Vext = inf within
    radius 3 of the origin, periodic) in a 12.8^3 cell, 128^3 grid.
c3_<i>: instance i of the 64-grid batch (deft_amd.workloads.c3_batch(64, seed=12345)):
    uneven edges from {32, 48, 64, 96}, dx = 0.1, eight Gaussians plus a planar hard
    wall, eta ~ U(0.1, 0.45) -- with the FP64 potential exactly as generated.
Both: mu from the bulk equation of state, Veff = -kT ln n_b start, new-API
preconditioned CG with relative precision 1e-12.  Each JSON holds the energy,
per-term energies, iteration counts, the density along the x axis through the
origin and a few whole-field checksums.
"""

import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fmt_oracle as O  # noqa: E402

CASES = {
    "c1": dict(n=64, L=6.4, eta=0.3, kind="wall"),
    "c2": dict(n=128, L=12.8, eta=0.4, kind="solute", radius=3.0),
}
# C3 batch instances with goldens: a 96 edge with mixed edges, the densest fluid, a cube
C3_GOLDEN = (12, 6, 1)
for _i in C3_GOLDEN:
    CASES["c3_%d" % _i] = dict(kind="c3", index=_i, seed=12345, count=64)


def problem(case):
    c = CASES[case]
    if c["kind"] == "c3":
        import deft_amd.workloads as W
        shape, cell, V, eta = W.c3_batch(c["count"], seed=c["seed"])[c["index"]]
        g = O.Grid(shape, np.diag(cell))
        c.update(shape=list(shape), cell=list(cell), eta=eta)
        return g, O.make_fluid(g, eta, vext=V)
    g = O.Grid.cubic(c["n"], c["L"])
    if c["kind"] == "wall":
        vext = O.planar_wall_vext(g, 1.0)
    else:
        vext = O.spherical_solute_vext(g, c["radius"])
    return g, O.make_fluid(g, c["eta"], vext=vext)


def run(case):
    g, f = problem(case)
    t0 = time.time()
    m = O.minimize(f, relative_precision=1e-12)
    wall = time.time() - t0
    E = m.energy()
    n = f.density()
    out = {
        "case": case, **CASES[case],
        "energy": E, "terms": f.last_terms,
        "iterations": m.iter, "energy_calcs": m.num_energy_calcs, "grad_calcs": m.num_grad_calcs,
        "stop_reason": m.stop_reason, "cpu_seconds": wall,
        "mu": f.mu,
        "density_x_axis": n[:, 0, 0].tolist(),
        "density_sum": float(n.sum()), "density_sumsq": float((n * n).sum()), "density_max": float(n.max()),
    }
    with open(os.path.join(HERE, f"oracle_{case}.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(case, E, m.iter, m.stop_reason, f"{wall:.1f}s", flush=True)


if __name__ == "__main__":
    for c in (sys.argv[1:] or list(CASES)):
        run(c)
