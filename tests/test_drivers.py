"""Driver output formats (deft_amd.drivers): the number format of src/handymath.h:51-68
against the C library's printf, and the radial/z plot layouts read back."""

import ctypes
import ctypes.util
import math

import numpy as np
import pytest

from deft_amd import drivers

_libc = ctypes.CDLL(ctypes.util.find_library("c"))


def _c_printf(fmt, width, prec, x):
    buf = ctypes.create_string_buffer(256)
    _libc.snprintf(buf, 256, fmt.encode(), ctypes.c_int(width), ctypes.c_int(prec), ctypes.c_double(x))
    return buf.value.decode()


@pytest.mark.parametrize("x", [0.0, 2.72225468848892, -5282.51062579, 1.5e-7, -3e-6, 1.234e7, -9.9e5, 1e-5, 0.5])
def test_print_double_matches_c(x):
    s = drivers.print_double("", x)
    body = s[1:]  # after the prefix and its space
    fixed = "." in body and "e" not in body
    ref = _c_printf("%*.*f" if fixed else "%*.*e", 26, 17 if fixed else 14, x)
    assert body == ref, (x, body, ref)
    # the window: fixed notation for moderate magnitudes (and zero), scientific outside
    assert fixed == (x == 0.0 or 1e-5 < abs(x) < 1e6), x


@pytest.mark.gpu
def test_radial_and_z_plots_read_back(gpu, tmp_path):
    import deft_amd.workloads as W
    shape, L = (32, 32, 32), 6.4
    nb = 0.3 / (4 * math.pi / 3)
    f = gpu.WhiteBearFluidVeff(shape, np.diag([L] * 3), mu=gpu.find_chemical_potential(nb), fft_precision=1)
    f.Vext = W.sphere_vext(shape, (L, L, L), (0.0, 0.0, 0.0), 1.5)
    f.fill_Veff(-math.log(nb))
    m = gpu.Minimize(f)
    m.precondition(True)
    m.set_relative_precision(1e-10)
    m.run()
    cv = f.contact_values()
    n = f.density()
    p = tmp_path / "inner.dat"
    drivers.radial_plot(str(p), f)
    rows = np.loadtxt(str(p))
    assert rows.shape == (16, 8)
    assert np.allclose(rows[:, 0], np.arange(16) * L / 32, rtol=1e-5)
    assert np.allclose(rows[:, 1], n[0, :16, 0], rtol=1e-5, atol=1e-12)
    ok = np.isfinite(cv["gSigmaS"][0, :16, 0])
    assert np.allclose(rows[ok, 3], cv["gSigmaS"][0, :16, 0][ok], rtol=1e-5)
    q = tmp_path / "walls.dat"
    drivers.z_plot(str(q), f)
    rz = np.loadtxt(str(q))
    assert rz.shape == (16, 9) and np.allclose(rz[:, 1], rz[:, 2])
    f.energy()
    text = drivers.printme(f, "  ")
    assert len(text.splitlines()) == 6
    info = drivers.print_info(m, f, "  ")
    assert "==== Iteration" in info and "total energy =" in info
    m.close()
    f.close()


def test_printme_layout():
    class F:
        last_terms = {"kTphi1": 1.5, "ideal": -2.25e-7, "total": 0.0}
    lines = drivers.printme(F(), "> ").splitlines()
    w = 1 + len("total energy")
    assert lines[0] == "> " + "kTphi1".rjust(w) + " =" + drivers.print_double("", 1.5)
    assert lines[1].startswith("> " + "ideal".rjust(w) + " = ") and "e-07" in lines[1]
