"""The oracle's C restatement (oracle/fmt_cpu.c), which bench.py times as the CPU
baseline, pinned against the numpy oracle and the reference's known answer.

* one energy + gradient evaluation (the new-API preconditioner included) on grids that
  exercise every radix of its FFT (edges 10 to 50, powers of 2, 3 and 5), a skewed cell,
  hard walls and constrained points, against oracle/fmt_oracle.py at round-off;
* the known answer 2.72225468848892 of tests/new-hard-spheres.cpp:52-63 (uniform n = 0.1,
  R = 1, cell 2^3 on 20^3 points);
* a whole minimization of the solute workload's shape (32^3) against the numpy oracle's.
"""

import json
import math
import os
import struct
import subprocess

import numpy as np
import pytest

from oracle import fmt_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fc") / "fmt_cpu")
    subprocess.run(["gcc", "-O2", "-std=c99", "-Wall", "-Wextra", "-Werror", os.path.join(ROOT, "oracle", "fmt_cpu.c"),
                    "-o", out, "-lm"], check=True)
    return out


def _eval(exe, tmp_path, shape, cell, R, kT, mu, vext, x):
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        fh.write(b"FMTC")
        fh.write(struct.pack("<3i", *shape))
        fh.write(np.ascontiguousarray(cell, dtype="<f8").tobytes())
        fh.write(struct.pack("<3d", R, kT, mu))
        fh.write(np.ascontiguousarray(vext, dtype="<f8").tobytes())
        fh.write(np.ascontiguousarray(x, dtype="<f8").tobytes())
    subprocess.run([exe, "eval", str(inp), str(outp)], check=True)
    d = np.fromfile(outp, dtype="<f8")
    N = int(np.prod(shape))
    return d[0], d[1:7], d[7:7 + N].reshape(shape), d[7 + N:].reshape(shape)


@pytest.mark.parametrize("shape,cell", [((32, 48, 20), np.diag([3.1, 4.9, 2.3])),
                                        ((16, 16, 16), np.array([[2.0, 0, 0.1], [0.2, 1.9, 0], [0, 0.3, 3.4]])),
                                        ((24, 30, 36), np.diag([2.4, 3.0, 3.6])),
                                        ((20, 10, 50), np.diag([2.0, 1.0, 5.0]))])
def test_evaluation_matches_numpy_oracle(exe, tmp_path, shape, cell):
    g = O.Grid(shape, cell)
    vext = O.planar_wall_vext(g, 0.5)
    vext = np.where(np.isfinite(vext), 0.3 * np.sin(g.positions()[..., 1]), vext)
    vext[shape[0] // 2] = np.nan
    of = O.make_fluid(g, 0.35, vext=vext)
    of.x = of.x + 0.2 * np.random.default_rng(1).standard_normal(g.shape)
    E, terms, gr, pg = _eval(exe, tmp_path, shape, cell, of.R, of.kT, of.mu, vext, of.x)
    Eo, gro, pgo = of.energy_grad_and_precond()
    assert abs(E - Eo) <= 1e-13 * abs(Eo)
    for t, k in zip(terms, ("kTphi1", "kTphi2", "kTphi3", "ideal", "int_V_n", "minus_mu_N")):
        assert abs(t - of.last_terms[k]) <= 1e-13 * abs(Eo), k
    assert np.max(np.abs(gr - gro)) <= 1e-13 * np.max(np.abs(gro))
    assert np.max(np.abs(pg - pgo)) <= 1e-13 * np.max(np.abs(pgo))


def test_white_bear_known_answer(exe, tmp_path):
    shape = (20, 20, 20)
    E, terms, _, _ = _eval(exe, tmp_path, shape, np.diag([2.0] * 3), 1.0, 1.0, 0.0, np.zeros(shape),
                           np.full(shape, -math.log(0.1)))
    assert abs((terms[0] + terms[1] + terms[2]) / 2.72225468848892 - 1) < 1e-13


def test_minimization_matches_numpy_oracle(exe):
    g = O.Grid.cubic(32, 3.2)
    of = O.make_fluid(g, 0.3, vext=O.spherical_solute_vext(g, 1.0))
    m = O.Minimize(of)
    m.precondition(True)
    m.set_relative_precision(1e-12)
    while m.improve_energy():
        pass
    r = json.loads(subprocess.run([exe, "solute", "32", "3.2", "0.3", "1.0", "0", "0", "0"], capture_output=True,
                                  text=True, check=True).stdout)
    assert r["finished"] and r["stop"] == "converged"
    assert abs(r["energy"] / m.energy() - 1) < 1e-12
