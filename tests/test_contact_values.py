"""Contact-density estimators (deft_contact_values) on the HIP path against the oracle:
n0, gSigmaS, YuWu_S, nA, gSigmaA, Gross and the C++-form gSigmaS of a walled,
solvated density, FP64, for White-Bear and both Mark II forms."""

import math

import numpy as np
import pytest

from oracle import fmt_oracle as O

pytestmark = pytest.mark.gpu

KEYS = ("n0", "gSigmaS", "YuWu_S", "nA", "gSigmaA", "Gross", "gSigmaS_cpp")
FORMS = [("wb", 0), ("wb2", 1), ("wb2slow", 2)]


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("shape,cell", [((32, 32, 32), np.diag([4.8, 4.8, 4.8])),
                                        ((21, 32, 20), np.diag([3.2, 4.8, 3.0]))])
def test_contact_values_match_oracle(gpu, shape, cell, form):
    print("contact", shape, form, flush=True)
    g = O.Grid(shape, cell)
    vext = O.spherical_solute_vext(g, 1.2, center=(1.0, 1.0, 1.0))
    of = O.make_fluid(g, 0.3, vext=vext, functional=form[0])
    of.x += 0.2 * np.random.default_rng(4).standard_normal(g.shape)
    f = gpu.WhiteBearFluidVeff(shape, cell, R=of.R, kT=of.kT, mu=of.mu, fft_precision=0, functional=form[1])
    f.Vext = of.vext
    f.Veff = of.x
    got = f.contact_values()
    ref = O.contact_values(of)
    n2 = of.weighted_densities(of.density())[1]
    ok = n2 > 1e-3 * n2.max()  # the estimators divide by n2 (and n0^2)
    for k, r in zip(KEYS, ref):
        a = got[k][ok]
        b = r[ok]
        assert np.all(np.isfinite(a)), k
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b)), (k, np.max(np.abs(a - b)))


def test_contact_values_of_a_uniform_fluid(gpu):
    eta = 0.4
    nb = O.bulk_density(eta, 1.0)
    f = gpu.WhiteBearFluidVeff((32, 32, 32), np.diag([3.2] * 3), mu=gpu.find_chemical_potential(nb), fft_precision=1)
    f.Veff = -math.log(nb)
    cv = f.contact_values()
    cs = (1 - eta / 2) / (1 - eta) ** 3
    for k in ("gSigmaS", "YuWu_S", "gSigmaA", "Gross", "gSigmaS_cpp"):
        assert np.max(np.abs(cv[k] / cs - 1)) < 1e-11, k
    assert np.max(np.abs(cv["nA"] / nb - 1)) < 1e-11
    f.functional = gpu.WHITE_BEAR_TENSOR
    with pytest.raises(gpu.DeftError):
        f.contact_values()
