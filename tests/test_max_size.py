"""The largest supported grid (256^3, 16.8 M points per field) through size-independent
properties: a uniform fluid is the bulk, the FFT seam round-trips, energies do not
depend on where a solute sits (whole-grid translations), and the energy-only and
energy+gradient evaluations agree bit for bit."""

import math

import numpy as np
import pytest

import deft_amd.workloads as W

pytestmark = pytest.mark.gpu

N, L = 256, 25.6


def test_uniform_fluid_at_256(gpu):
    nb = 0.4 / (4 * math.pi / 3)
    f = gpu.WhiteBearFluidVeff((N, N, N), np.diag([L] * 3), mu=gpu.find_chemical_potential(nb), fft_precision=1)
    f.fill_Veff(-math.log(nb))
    E, gr, pg = f.energy_grad_and_precond()
    p = gpu.pressure(nb)
    assert abs(E + p * L ** 3) < 1e-11 * p * L ** 3
    assert np.max(np.abs(pg)) < 1e-11
    f.close()


@pytest.mark.parametrize("prec,tol", [(1, 1e-13), (0, 3e-6)])
def test_fft_round_trip_at_256(gpu, prec, tol):
    f = gpu.WhiteBearFluidVeff((N, N, N), np.diag([L] * 3), fft_precision=prec)
    a = np.random.default_rng(1).standard_normal((N, N, N))
    back = f.ifft(f.fft(a))
    assert np.max(np.abs(back - a)) < tol * np.max(np.abs(a))
    f.close()


@pytest.mark.parametrize("prec,tol", [(1, 1e-12), (0, 1e-6)])
def test_translation_invariance_at_256(gpu, prec, tol):
    nb = 0.35 / (4 * math.pi / 3)
    shift = (17, 5, 100)
    V0 = W.sphere_vext((N, N, N), (L, L, L), (0.0, 0.0, 0.0), 3.0)
    bump = 0.1 * np.cos(2 * math.pi * np.arange(N) / N)[:, None, None] * np.ones((1, N, N))
    X0 = np.where(np.isfinite(V0), -math.log(nb), 0.0) + bump
    out = []
    for s in ((0, 0, 0), shift):
        # the translated state is an exact roll of the first (no re-sampled boundary points)
        f = gpu.WhiteBearFluidVeff((N, N, N), np.diag([L] * 3), mu=gpu.find_chemical_potential(nb), fft_precision=prec)
        f.Vext = np.roll(V0, s, axis=(0, 1, 2))
        f.Veff = np.roll(X0, s, axis=(0, 1, 2))
        e = f.energy()
        E, gr, _ = f.energy_grad_and_precond()
        assert e == E
        out.append((E, gr))
        f.close()
    (E0, g0), (E1, g1) = out
    assert abs(E1 - E0) <= tol * abs(E0)
    g0s = np.roll(g0, shift, axis=(0, 1, 2))
    assert np.max(np.abs(g1 - g0s)) <= 10 * tol * np.max(np.abs(g0))
