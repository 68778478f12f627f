"""deft_functional_grad_noalloc against the entry point it varies,
deft_functional_grad_device_kind, through the C ABI on device buffers.

Both run the same passes; the allocation-free variant takes the grid as
parameters and keeps every intermediate in the library's static device array.
Two calls in a row (the second on a smaller grid, so the array holds the first
call's data), both FFT precisions and both pgrad kinds; outputs accumulate (+=)
into buffers that hold data beforehand.  Bar: FP64 energy within 1e-13 and
fields within 1e-12 of the field scale (the twiddles are computed on the device
here, on the host in the plan: they may differ in the last bit); FP32 energy
within 1e-7 and fields within 1e-6.
"""

import ctypes
import math

import numpy as np
import pytest

from oracle import fmt_oracle as O

pytestmark = pytest.mark.gpu


def _buf(L, nbytes):
    p = ctypes.c_void_p()
    assert L.deft_buffer_create(nbytes, ctypes.byref(p)) == 0
    return p


def _up(L, p, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert L.deft_buffer_upload(p, 0, a.ctypes.data_as(ctypes.c_void_p), a.nbytes) == 0


def _down(L, p, n):
    out = np.empty(n, dtype=np.float64)
    assert L.deft_buffer_download(p, 0, out.ctypes.data_as(ctypes.c_void_p), out.nbytes) == 0
    return out


def _case(gpu, shape, cell, prec, kind, seed):
    L = gpu.lib()
    N = int(np.prod(shape))
    g = O.Grid(shape, cell)
    rng = np.random.default_rng(seed)
    # smooth potential FP32 cannot represent, a hard wall and a constrained slab
    vext = 0.7 * np.sin(2 * math.pi * g.positions()[..., 1] / cell[1][1]) + 1e-3 * rng.standard_normal(g.shape)
    vext[: max(1, shape[0] // 8)] = np.inf
    vext[shape[0] // 2] = np.nan
    of = O.make_fluid(g, 0.35, vext=vext)
    veff = of.x + 0.2 * rng.standard_normal(g.shape)
    f = gpu.WhiteBearFluidVeff(shape, cell, R=of.R, kT=of.kT, mu=of.mu, fft_precision=prec)
    f.Vext = vext
    bufs = [_buf(L, 8 * N) for _ in range(7)]
    b_veff, b_vext, b_g1, b_pg1, b_g2, b_pg2, b_e = bufs
    _up(L, b_veff, veff)
    _up(L, b_vext, vext)
    pre_g = rng.standard_normal(N)
    pre_pg = rng.standard_normal(N)
    for b, a in ((b_g1, pre_g), (b_g2, pre_g), (b_pg1, pre_pg), (b_pg2, pre_pg)):
        _up(L, b, a)
    E1 = ctypes.c_double()
    gpu._check(L.deft_functional_grad_device_kind(f.handle, ctypes.cast(b_veff, gpu._DP), ctypes.cast(b_g1, gpu._DP),
                                                   ctypes.cast(b_pg1, gpu._DP), ctypes.byref(E1), kind))
    gpu.functional_grad_noalloc(shape, cell, of.R, of.kT, of.mu, prec, kind, b_veff, b_vext, b_g2, b_pg2, b_e)
    E2 = _down(L, b_e, 1)[0]
    g1, g2 = _down(L, b_g1, N), _down(L, b_g2, N)
    p1, p2 = _down(L, b_pg1, N), _down(L, b_pg2, N)
    for b in bufs:
        L.deft_buffer_destroy(b)
    f.close()
    return E1.value, E2, g1 - pre_g, g2 - pre_g, p1 - pre_pg, p2 - pre_pg


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("prec,etol,ftol", [(1, 1e-13, 1e-12), (0, 1e-7, 1e-6)])
def test_noalloc_matches_the_plan_entry_point(gpu, prec, etol, ftol):
    # two calls in a row, the second on a smaller grid; both pgrad kinds
    cases = [((32, 48, 20), np.diag([3.1, 4.9, 2.3]), gpu.PGRAD_NEW, 1),
             ((16, 16, 16), np.diag([1.7, 1.6, 1.8]), gpu.PGRAD_EFFECTIVE_POTENTIAL, 2)]
    for shape, cell, kind, seed in cases:
        print("noalloc", shape, prec, kind, flush=True)
        E1, E2, g1, g2, p1, p2 = _case(gpu, shape, cell, prec, kind, seed)
        assert np.isfinite(E1) and abs(E2 - E1) <= etol * abs(E1), (E1, E2)
        assert _rel(g2, g1) < ftol
        assert _rel(p2, p1) < ftol


def test_noalloc_largest_grid_and_rejections(gpu):
    shape, cell = (64, 64, 64), np.diag([6.4, 6.4, 6.4])
    print("noalloc 64^3", flush=True)
    E1, E2, g1, g2, p1, p2 = _case(gpu, shape, cell, 1, gpu.PGRAD_NEW, 3)
    assert abs(E2 - E1) <= 1e-13 * abs(E1)
    assert _rel(g2, g1) < 1e-12 and _rel(p2, p1) < 1e-12
    L = gpu.lib()
    cell = np.ascontiguousarray(np.diag([9.6, 6.4, 6.4]))
    rc = L.deft_functional_grad_noalloc(96, 64, 64, cell.ctypes.data_as(gpu._DP), 1.0, 1.0, 0.0, 1, gpu.PGRAD_NEW,
                                        None, None, None, None, None, None)
    assert rc == 1  # DEFT_ERR_ARG: null buffers are checked first
    dummy = ctypes.c_void_p(1)
    rc = L.deft_functional_grad_noalloc(96, 64, 64, cell.ctypes.data_as(gpu._DP), 1.0, 1.0, 0.0, 1, gpu.PGRAD_NEW,
                                        dummy, None, dummy, None, None, None)
    assert rc == 1 and b"262144" in L.deft_last_error()
