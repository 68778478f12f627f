"""Synthetic workloads of the hard-sphere minimization path (R = 1, kT = 1).

  C1  64^3 periodic 6.4^3 cell, hard wall Vext = inf for x < 1, eta = 0.3
      (papers/contact/figs/walls.cpp:83-113 shape)
  C2  128^3 periodic 12.8^3 cell, hard solute Vext = inf within radius 3 of the
      origin, eta = 0.40 (papers/contact/figs/inner-sphere.cpp:177-211 shape;
      radius 3 is the excluded radius of sphere centres)
  C3  a batch of uneven periodic grids, edges per axis from {32, 48, 64, 96} at
      dx = 0.1, Vext = 8 Gaussians A exp(-|r-c|^2 / 2w^2), A ~ U(-2, 2) kT,
      w ~ U(0.5, 2), plus one planar hard wall of thickness R on a random axis,
      eta ~ U(0.1, 0.45)
  C5  256^3 periodic 12.8^3 cell (dx = 0.05), 30 hard spherical obstacles of
      radius ~ U(1, 2.5) (excluded radius of centres), eta = 0.45

Every generator is seeded; the same seed gives the same instances.
"""

import math

import numpy as np

C3_EDGES = (32, 48, 64, 96)
DX = 0.1


def bulk_density(eta, R=1.0):
    return eta / (4 * math.pi / 3 * R ** 3)


def _axis(n, length, c=0.0):
    """Minimum-image offsets of the grid coordinates i*length/n from c."""
    x = np.arange(n) * (length / n) - c
    return x - length * np.round(x / length)


def planar_wall_vext(shape, cell, thickness=1.0):
    """C1: Vext = +inf where x < thickness."""
    x = np.arange(shape[0]) * (cell[0] / shape[0])
    return np.broadcast_to(np.where(x < thickness, np.inf, 0.0)[:, None, None], shape).copy()


def sphere_vext(shape, cell, centre, radius, out=None):
    """Vext = +inf within `radius` (periodic minimum image) of `centre`."""
    V = np.zeros(shape) if out is None else out
    dx, dy, dz = (_axis(n, L, c) for n, L, c in zip(shape, cell, centre))
    r = np.sqrt(dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2)
    V[r < radius] = np.inf
    return V


def c2_solute_vext(n=128, length=12.8, radius=3.0):
    return sphere_vext((n, n, n), (length,) * 3, (0.0, 0.0, 0.0), radius)


def c3_vext(shape, cell, rng, R=1.0):
    """The C3 potential on a grid: 8 Gaussians A exp(-|r-c|^2 / 2w^2) with
    A ~ U(-2, 2) kT, w ~ U(0.5, 2), c uniform, plus one planar hard wall of
    thickness R on a random axis at a random offset."""
    V = np.zeros(shape)
    for _ in range(8):
        A = rng.uniform(-2.0, 2.0)
        w = rng.uniform(0.5, 2.0)
        c = [rng.uniform(0, L) for L in cell]
        dx, dy, dz = (_axis(n, L, ci) for n, L, ci in zip(shape, cell, c))
        V += A * np.exp(-(dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2) / (2 * w * w))
    axis = int(rng.integers(0, 3))
    wall = np.abs(_axis(shape[axis], cell[axis], rng.uniform(0, cell[axis]))) < R / 2
    sl = [None, None, None]
    sl[axis] = slice(None)
    return np.where(np.broadcast_to(wall[tuple(sl)], shape), np.inf, V)


def c3_instance(rng, edges=C3_EDGES, R=1.0):
    """One C3 instance: (shape, cell, vext, eta); edges drawn per axis from `edges`."""
    shape = tuple(int(rng.choice(edges)) for _ in range(3))
    cell = tuple(n / 10.0 for n in shape)  # dx = 0.1
    V = c3_vext(shape, cell, rng, R)
    eta = float(rng.uniform(0.1, 0.45))
    return shape, cell, V, eta


def c3_cubic_vext(M, L, rng, R=1.0):
    """The C3 potential on a cubic M^3 grid of edge L."""
    return c3_vext((M, M, M), (L, L, L), rng, R)


def c3_batch(count=64, seed=12345, edges=C3_EDGES):
    rng = np.random.default_rng(seed)
    return [c3_instance(rng, edges) for _ in range(count)]


def c5_instance(rng, n=256, length=12.8, obstacles=30):
    """One C5 instance: (shape, cell, vext, eta = 0.45)."""
    shape, cell = (n, n, n), (length,) * 3
    V = np.zeros(shape)
    for _ in range(obstacles):
        sphere_vext(shape, cell, [rng.uniform(0, length) for _ in range(3)], rng.uniform(1.0, 2.5), out=V)
    return shape, cell, V, 0.45


def c5_batch(count=16, seed=2024, n=256):
    rng = np.random.default_rng(seed)
    return [c5_instance(rng, n=n) for _ in range(count)]
