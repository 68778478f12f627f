"""The library's minimizer control logic (deft_amd/csrc/minimizer.hpp) on the CPU.

tests/host/minimizer_harness.cpp compiles the same header with g++ against a
host backend on the quadratic E(x) = x.A.x/2 - b.x.  Checked here:

* the old-API stack (ConjugateGradient / PreconditionedConjugateGradient with
  QuadraticLineMinimizer, src/ConjugateGradient.cpp:59-160,
  src/QuadraticLineMinimizer.cpp:27-198) against the oracle's restatement,
  iteration by iteration, and both against mathematics: the quadratic line
  search is exact on a quadratic, so Polak-Ribiere CG is linear CG and reaches
  the minimum of an n-dimensional quadratic in n iterations (a wrong curvature
  sign, slope convention or beta fails this);
* Precision (src/Precision.cpp:23-156) around it, against the oracle;
* the new API's Minimize (src/new/Minimize.cpp:9-313) against the oracle;
* the precision ladder's final verdict: with FP64 energies that carry a small
  roughness (a stand-in for the round-off of FFT-based energies) the
  energy-based estimate alone reports convergence well above the target (the
  false stop); the gradient confirmation turns every such stop into an
  unconverged one and raises no false alarm on smooth problems.
"""

import math
import os
import subprocess

import numpy as np
import pytest

from oracle import fmt_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS_SRC = os.path.join(ROOT, "tests", "host", "minimizer_harness.cpp")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("mh") / "minimizer_harness")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                    HARNESS_SRC, "-o", exe], check=True)
    return exe


def _fmt(a):
    return " ".join(format(float(v), ".17g") for v in np.ravel(a))


class Quadratic:
    """E = x.A.x/2 - b.x with the oracle minimizers' interface."""

    def __init__(self, A, b, x0, dinv):
        self.A, self.b, self.x, self.dinv = A, b, np.array(x0, dtype=np.float64), dinv

    def energy(self):
        return 0.5 * float(self.x @ (self.A @ self.x)) - float(self.b @ self.x)

    def grad(self):
        return self.A @ self.x - self.b

    def grad_and_pgrad(self):
        g = self.grad()
        return g, self.dinv * g

    def energy_grad_and_precond(self):
        g = self.grad()
        return self.energy(), g, self.dinv * g

    def have_preconditioner(self):
        return True


def _problem(seed, n, mmax, rotate_only=False):
    """A = D^1/2 M D^1/2: M has eigenvalues geometrically in [1, mmax] (random basis), D is
    a diagonal spread over two decades; dinv = 1/diag(D) -- diagonally dominated like the
    fluid, whose ideal-gas Hessian is diagonal."""
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    M = Q @ np.diag(np.geomspace(1, mmax, n)) @ Q.T
    dh = np.sqrt(np.geomspace(0.1, 10, n)[rng.permutation(n)])
    A = dh[:, None] * M * dh[None, :]
    A = 0.5 * (A + A.T)
    b = rng.normal(size=n) * 10
    x0 = np.zeros(n)
    dinv = 1 / dh ** 2
    xs = np.linalg.solve(A, b)
    Emin = 0.5 * xs @ A @ xs - b @ xs
    return A, b, x0, dinv, Emin


def _run(exe, args, A, b, x0, dinv):
    inp = f"{len(b)}\n{_fmt(A)}\n{_fmt(b)}\n{_fmt(x0)}\n{_fmt(dinv)}\n"
    r = subprocess.run([exe] + [str(a) for a in args], input=inp, capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    return [l.split() for l in lines[:-1]], lines[-1].split()


@pytest.mark.parametrize("preconditioned", [False, True])
def test_old_cg_is_linear_cg_on_a_quadratic(harness, preconditioned):
    n = 8
    A, b, x0, dinv, Emin = _problem(1, n, 30.0)
    # the oracle: exact minimum after n iterations
    q = Quadratic(A, b, x0, dinv)
    cg = O.OldConjugateGradient(q, preconditioned=preconditioned, stepsize=10.0)
    oracle_E = []
    for _ in range(n):
        cg.improve_energy()
        oracle_E.append(q.energy())
    assert abs(oracle_E[-1] - Emin) <= 1e-12 * abs(Emin)
    assert oracle_E[-2] - Emin > 1e-9 * abs(Emin)  # not earlier: the n-step bound is sharp here
    # the library's logic, iteration by iteration
    rows, final = _run(harness, ["old", int(preconditioned), 10.0, 0.0, n], A, b, x0, dinv)
    assert len(rows) == n
    for k, row in enumerate(rows):
        assert int(row[0]) == k + 1
        assert abs(float(row[1]) - oracle_E[k]) <= 1e-12 * abs(Emin), k
    assert abs(float(final[2]) - Emin) <= 1e-12 * abs(Emin)


def test_old_cg_ill_conditioned_agrees_with_oracle(harness):
    """An ill-conditioned quadratic (condition ~3e4): library and oracle take the same
    branches and steps; plain CG amplifies round-off there, so only the first 12 iterations
    are compared (later ones drift apart at the 1e-3 level without any branch differing)."""
    A, b, x0, dinv, Emin = _problem(4, 30, 300.0)
    q = Quadratic(A, b, x0, dinv)
    cg = O.OldConjugateGradient(q, preconditioned=False, stepsize=10.0)
    rows, _ = _run(harness, ["old", 0, 10.0, 0.0, 12], A, b, x0, dinv)
    assert len(rows) == 12
    for row in rows:
        cg.improve_energy()
        assert abs(float(row[1]) - q.energy()) <= 1e-11 * abs(Emin)
        assert float(row[2]) == pytest.approx(cg.step, rel=1e-9)


@pytest.mark.parametrize("preconditioned,tol", [(False, 1e-6), (True, 1e-12)])
def test_precision_wrapper_matches_oracle(harness, preconditioned, tol):
    """Same stop and energies.  Plain CG works in the unwhitened coordinates (condition
    ~1e3 here) and amplifies round-off between the two implementations to ~1e-7 of E;
    the preconditioned form stays at round-off."""
    A, b, x0, dinv, Emin = _problem(5, 16, 10.0)
    precision = 1e-8 * abs(Emin)
    q = Quadratic(A, b, x0, dinv)
    m = O.OldPrecision(precision, O.OldConjugateGradient(q, preconditioned=preconditioned, stepsize=10.0))
    oracle = []
    while True:
        more = m.improve_energy()
        oracle.append(q.energy())
        if not more or len(oracle) > 500:
            break
    rows, final = _run(harness, ["old", int(preconditioned), 10.0, precision, 500], A, b, x0, dinv)
    assert len(rows) == len(oracle)
    for row, Eo in zip(rows, oracle):
        assert abs(float(row[1]) - Eo) <= tol * abs(Emin)
    assert int(final[1]) == 9  # DEFT_STOP_CONVERGED
    # the estimate is heuristic; the answer is within a few times the precision
    assert float(final[2]) - Emin < 10 * precision


@pytest.mark.parametrize("preconditioned", [False, True])
def test_new_minimize_matches_oracle(harness, preconditioned):
    """Preconditioned: every iterate at round-off and the same stop.  Plain: the first ten
    iterates (round-off then spreads, see above) and both answers at the minimum."""
    A, b, x0, dinv, Emin = _problem(3, 30, 30.0)
    q = Quadratic(A, b, x0, dinv)
    m = O.Minimize(q)
    m.precondition(preconditioned)
    m.set_relative_precision(1e-12)
    oracle = []
    while True:
        more = m.improve_energy()
        oracle.append(m.energy())
        if not more or len(oracle) > 400:
            break
    rows, final = _run(harness, ["new", int(preconditioned), 1e-12, 400], A, b, x0, dinv)
    if preconditioned:
        assert len(rows) == len(oracle)
    for row, Eo in list(zip(rows, oracle))[:None if preconditioned else 10]:
        assert abs(float(row[1]) - Eo) <= 1e-11 * abs(Emin)
    assert abs(float(final[2]) - Emin) < 1e-9 * abs(Emin)
    assert abs(oracle[-1] - Emin) < 1e-9 * abs(Emin)


def _ladder(exe, seed, n, mmax, noise, grad_check, rel=1e-10):
    A, b, x0, dinv, Emin = _problem(seed, n, mmax)
    precision = rel * abs(Emin)
    rows, final = _run(exe, ["ladder", 1e-5, noise, grad_check, format(precision, ".17g"), 600], A, b, x0, dinv)
    return int(final[1]), (float(final[2]) - Emin) / precision


@pytest.mark.parametrize("seed,n,mmax", [(0, 20, 10.0), (1, 20, 30.0), (2, 20, 30.0), (3, 20, 10.0)])
def test_ladder_false_stop_is_caught_by_the_gradient(harness, seed, n, mmax):
    # energy-only verdict (confirmation off): "converged" at more than 10x the target
    stop0, err0 = _ladder(harness, seed, n, mmax, 1e-9, 0.0)
    assert stop0 == 9 and err0 > 10
    # with the gradient confirmation the same run does not claim convergence
    stop, err = _ladder(harness, seed, n, mmax, 1e-9, 10.0)
    assert stop != 9
    assert err > 10


@pytest.mark.parametrize("seed,n,mmax,noise", [(s, n, m, z) for s in range(4) for n in (20, 40)
                                               for m in (10.0, 30.0) for z in (0.0, 1e-12)])
def test_ladder_gradient_check_raises_no_false_alarm(harness, seed, n, mmax, noise):
    stop, err = _ladder(harness, seed, n, mmax, noise, 10.0)
    assert stop == 9
    assert err < 1.0


@pytest.mark.parametrize("seed,n", [(s, n) for s in range(4) for n in (20, 40)])
@pytest.mark.parametrize("amp", [1e-9, 1e-8, 1e-7])
def test_ladder_start_near_the_minimum_converges(harness, seed, n, amp):
    """A start within 1e-9 - 1e-7 of the minimum: the FP32 phase hands over at once and
    the all-FP64 phase's line searches find energies that rise or stay at round-off
    level before any iteration has met the precision.  The gradient at that point
    decides and confirms convergence; a verdict from the energy alone ends most of
    these runs as 'energy-rose' or 'no-change' although they sit at the minimum."""
    A, b, _, dinv, Emin = _problem(seed, n, 10.0)
    xs = np.linalg.solve(A, b)
    x0 = xs + amp * np.random.default_rng(seed).normal(size=n)
    precision = 1e-10 * abs(Emin)
    rows, final = _run(harness, ["ladder", 1e-5, 0.0, 10.0, format(precision, ".17g"), 600], A, b, x0, dinv)
    assert int(final[1]) == 9
    assert abs(float(final[2]) - Emin) < precision
