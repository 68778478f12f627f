"""Output formats of the reference's drivers, written from a HIP-path fluid, so their
plot scripts read our results unchanged.

* ``radial_plot``: papers/contact/figs/inner-sphere.cpp:143-172 — for y = 0..Ny/2-1 the
  line "y  n  n0  gSigmaS  YuWu_S  nA  gSigmaA  Gross" at grid point (0, y, 0),
  tab-separated, "%g" (the driver's call at :229-231 fixes the column order).
* ``z_plot``: papers/contact/figs/walls.cpp:55-80 — for z = 0..Nz/2-1 at (0, 0, z):
  "z  n  n  gSigmaS  YuWu_S  gSigmaA  n0  Gross  nA" (the call at walls.cpp:128-129).
* ``print_double`` / ``printme``: the generated new-API ``printme`` (one line per named
  energy term, src/haskell/NewCode.hs:244-258) with src/handymath.h:51-68's number format.
"""

import numpy as np

from . import WHITE_BEAR_TENSOR


def _columns(fluid):
    if fluid.functional == WHITE_BEAR_TENSOR:
        raise ValueError("contact estimators are defined for White-Bear and Mark II")
    n = np.asarray(fluid.density())
    cv = fluid.contact_values()
    return n, cv


def _g(x):
    return "%g" % x


def radial_plot(path, fluid):
    """inner-sphere.cpp's radial_plot along y through the origin (a cell with the
    lattice vector a2 along y, as the driver's cubic cell)."""
    n, cv = _columns(fluid)
    Ny = fluid.shape[1]
    a2 = np.asarray(fluid.cell, dtype=np.float64).reshape(3, 3)[1]
    with open(path, "w") as fh:
        for y in range(Ny // 2):
            here = a2 * (y / Ny)
            row = [here[1], n[0, y, 0]] + [cv[k][0, y, 0] for k in ("n0", "gSigmaS", "YuWu_S", "nA", "gSigmaA", "Gross")]
            fh.write("\t".join(_g(v) for v in row) + "\n")


def z_plot(path, fluid):
    """walls.cpp's z_plot along z through (0, 0)."""
    n, cv = _columns(fluid)
    Nz = fluid.shape[2]
    a3 = np.asarray(fluid.cell, dtype=np.float64).reshape(3, 3)[2]
    with open(path, "w") as fh:
        for z in range(Nz // 2):
            here = a3 * (z / Nz)
            row = [here[2], n[0, 0, z], n[0, 0, z]] + [cv[k][0, 0, z] for k in ("gSigmaS", "YuWu_S", "gSigmaA", "n0",
                                                                               "Gross", "nA")]
            fh.write("\t".join(_g(v) for v in row) + "\n")


def print_double(prefix, x, width=26, digits=14, min_digits=11):
    """The string handymath.h's print_double prints: fixed notation inside a window
    around 1 set by width and digits, scientific outside it."""
    if min_digits == 0 or min_digits > digits:
        min_digits = digits
    width -= len(prefix)
    if width < digits + 8:
        width = digits + 8
    if x > 0:
        max_f = min(10.0 ** (width - digits - 3), 1e6)
        min_f = max(10.0 ** (min_digits - digits - 4), 1e-5)
    else:
        min_f = max(-(10.0 ** (width - digits - 5)), -1e6)
        max_f = min(-(10.0 ** (min_digits - digits - 4)), -1e-5)
    if (min_f < x < max_f) or x == 0.0:
        return "%s %*.*f" % (prefix, width, digits + 3, x)
    return "%s %*.*e" % (prefix, width, digits, x)


def printme(fluid, prefix=""):
    """The per-term energy report of the generated printme for the last evaluation
    (names padded to one more than the longest, "total energy" included in that
    width, then print_double of the value), one line per term."""
    terms = dict(fluid.last_terms)
    if not terms:
        raise ValueError("no evaluation yet")
    terms.pop("total", None)
    width = 1 + max(len(s) for s in list(terms) + ["total energy"])
    return "\n".join("%s%s =%s" % (prefix, name.rjust(width), print_double("", v)) for name, v in terms.items())


def print_info(minimizer, fluid, prefix="", with_iteration=True):
    """Minimize::print_info (src/new/Minimize.cpp:315-321): the iteration header, printme
    and the total energy."""
    out = ""
    if with_iteration:
        out += "\n%s==== Iteration %d ====\n" % (prefix, minimizer.stats()["iter"])
    out += printme(fluid, prefix) + "\n"
    out += "%s total energy =%s\n\n" % (prefix, print_double("", minimizer.energy()))
    return out
