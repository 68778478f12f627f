"""The reference's own converged solute profiles, reproduced through the HIP path.

Fixtures (tests/golden/contact_inner_sphere.json, written by make_contact_fixtures.py)
are papers/contact/figs/inner-sphere{WB,WBT,WBm2}-04.0-{eta}.dat: a hard solute of
excluded-centre radius 2 in a periodic 20^3 cell on a 201^3 grid, R = kT = 1, mu from
the bulk of the same functional, Veff started at -ln(1.0001 n_b) outside
(papers/contact/figs/inner-sphere.cpp:177-249); the density along y through the
centre, printed to 6 significant digits.

Bar: outside the solute every point within 4e-4 (relative) and the mean within 5e-5.
The estimator columns (n0, nA, YuWu_S, Gross) of all ten files within 1e-4.  The
gSigma columns of the Mark II files at eta 0.10 and 0.40 (computed with the
Haskell, 24 pi, Mark II Phi on the 36 pi profile) within 3e-5: gSigmaA, and
gSigmaS in the C++ by-hand form (src/ContactDensity.cpp:269-355), which those
files follow instead of today's Haskell helper.  The other files' gSigma columns
follow neither form of today's source (up to 0.13 off near the solute), so they
are printed, not pinned.
The reference's runs stopped at |E_bulk V| 1e-10 after at most 200 old-API iterations
and the profile is printed with %g, which set the floor; WB agrees to 5e-5 at worst.
The Mark II files match HardSpheresWBm2slow's form (Phi3 over 36 pi,
src/HardSpheres.cpp:161-185); the Haskell form WhiteBearMarkII, which today's
driver would generate (src/haskell/WhiteBear.hs:140-159), misses them by about 4 %
at contact, as a test below records.

The published tensor White-Bear files (inner-sphereWBT-04.0-{0.30,0.40}.dat) coincide
with the Mark I files to 2.3e-5, so the tensor term had no visible effect in the runs
that wrote them; the tensor functional as the reference source defines it
(src/haskell/WhiteBear.hs:44-51 with src/haskell/FMT.hs:82-88, pinned by the oracle
tests) moves this profile by up to about 1e-3.  The last test records both facts:
Mark I meets those files at the bar, the tensor functional sits about 1e-3 away.
"""

import json
import math
import os

import numpy as np
import pytest

import deft_amd
import deft_amd.workloads as W

pytestmark = pytest.mark.gpu

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "contact_inner_sphere.json")))
KEYS = sorted(k for k in FIX["profiles"] if not k.startswith("WBT"))
# published files whose gSigma columns today's formulas reproduce: the Mark II runs
# at eta 0.10 and 0.40, the ones regenerated with the averaged "-mean" companions
# (the other files came from older versions of the estimators)
GSIGMA_PINNED = ("WBm2-0.10", "WBm2-0.40")
TENSOR_KEYS = sorted(k for k in FIX["profiles"] if k.startswith("WBT"))


@pytest.fixture(scope="module")
def solute():
    n, L = FIX["n"], FIX["cell"]
    return W.sphere_vext((n, n, n), (L, L, L), (0.0, 0.0, 0.0), FIX["diameter"] / 2)


def _profile(V, eta, functional, contact=None):
    n, L = FIX["n"], FIX["cell"]
    nb = eta / (4 * math.pi / 3)
    f = deft_amd.WhiteBearFluidVeff((n, n, n), np.diag([L] * 3), R=FIX["R"], kT=FIX["kT"],
                                    mu=deft_amd.find_chemical_potential(nb, functional=functional),
                                    fft_precision=deft_amd.FFT_FP32, functional=functional)
    f.Vext = V
    f.fill_Veff(-math.log(nb))
    m = deft_amd.Minimize(f)
    m.precondition(True)
    m.set_relative_precision(1e-12)
    m.set_mixed_precision(1e-5, deft_amd.FFT_FP64)
    m.set_maxiter(3000)
    m.run()
    s = m.stats()
    prof = np.asarray(f.density())[0, :n // 2, 0].copy()
    cv = None
    if contact is not None:
        f.functional = contact
        cv = {k: v[0, :n // 2, 0].copy() for k, v in f.contact_values().items()}
    m.close()
    f.close()
    if contact is not None:
        return s, prof, cv
    return s, prof


@pytest.mark.parametrize("key", KEYS)
def test_inner_sphere_profile_matches_reference(gpu, solute, key):
    fx = FIX["profiles"][key]
    mark2 = key.startswith("WBm2")
    functional = deft_amd.WHITE_BEAR_MARK_II_SLOW if mark2 else deft_amd.WHITE_BEAR
    print("profile", key, flush=True)
    # the published Mark II estimator columns use the Haskell (24 pi) Phi
    s, prof, cv = _profile(solute, fx["eta"], functional,
                           contact=deft_amd.WHITE_BEAR_MARK_II if mark2 else deft_amd.WHITE_BEAR)
    assert s["stop_reason"] == "converged", s
    y = np.array(fx["y"])
    ref = np.array(fx["density"])
    assert np.allclose(y, np.arange(len(y)) * FIX["cell"] / FIX["n"], rtol=1e-5, atol=1e-6)
    out = y > FIX["diameter"] / 2
    rel = np.abs(prof[out] - ref[out]) / ref[out]
    print(key, "iterations", s["iter"], "max rel", rel.max(), "mean rel", rel.mean(), flush=True)
    assert rel.max() < 4e-4 and rel.mean() < 5e-5
    assert np.all(prof[~out] == 0)
    # contact-density estimators written beside the profile (inner-sphere.cpp:220-233)
    for c in ("n0", "nA", "YuWu_S", "Gross"):
        ref_c = np.array(fx[c])
        rel_c = np.abs(cv[c][out] - ref_c[out]) / np.abs(ref_c[out])
        print(key, c, "max rel", rel_c.max(), flush=True)
        assert rel_c.max() < 1e-4, c
    for c in ("gSigmaA", "gSigmaS", "gSigmaS_cpp"):
        ref_c = np.array(fx["gSigmaS" if c == "gSigmaS_cpp" else c])
        rel_c = np.abs(cv[c][out] - ref_c[out]) / np.abs(ref_c[out])
        pinned = key in GSIGMA_PINNED and c != "gSigmaS"
        print(key, c, "max rel", rel_c.max(), "(pinned)" if pinned else "(not pinned)", flush=True)
        if pinned:
            assert rel_c.max() < 3e-5, c


@pytest.mark.parametrize("key", ["WB-0.40", "WBm2-0.10"])
def test_reference_constraint_setup_reproduces_the_whole_column(gpu, key):
    """The driver's own set-up (inner-sphere.cpp:186-198): constrain(mask, f) with the
    solute points masked (NaN in Vext) and Veff started at -ln(n_b mask + 1e-4 n_b);
    the frozen interior then shows the 1e-4 n_b density the files print there."""
    fx = FIX["profiles"][key]
    n, L = FIX["n"], FIX["cell"]
    eta = fx["eta"]
    nb = eta / (4 * math.pi / 3)
    functional = deft_amd.WHITE_BEAR_MARK_II_SLOW if key.startswith("WBm2") else deft_amd.WHITE_BEAR
    wall = W.sphere_vext((n, n, n), (L, L, L), (0.0, 0.0, 0.0), FIX["diameter"] / 2)
    inside = np.isinf(wall)
    f = deft_amd.WhiteBearFluidVeff((n, n, n), np.diag([L] * 3), R=FIX["R"], kT=FIX["kT"],
                                    mu=deft_amd.find_chemical_potential(nb, functional=functional),
                                    fft_precision=deft_amd.FFT_FP32, functional=functional)
    f.Vext = np.where(inside, np.nan, 0.0)
    f.Veff = -np.log(np.where(inside, 0.0, nb) + 1e-4 * nb)
    m = deft_amd.Minimize(f)
    m.precondition(True)
    m.set_relative_precision(1e-12)
    m.set_mixed_precision(1e-5, deft_amd.FFT_FP64)
    m.set_maxiter(3000)
    m.run()
    assert m.stats()["stop_reason"] == "converged"
    prof = np.asarray(f.density())[0, :n // 2, 0]
    ref = np.array(fx["density"])
    rel = np.abs(prof - ref) / ref
    print(key, "whole column max rel", rel.max(), "mean", rel.mean(), flush=True)
    assert rel.max() < 4e-4 and rel.mean() < 5e-5
    m.close()
    f.close()


def test_haskell_mark_ii_is_not_what_the_published_profiles_used(gpu, solute):
    fx = FIX["profiles"]["WBm2-0.40"]
    _, prof = _profile(solute, fx["eta"], deft_amd.WHITE_BEAR_MARK_II)
    y = np.array(fx["y"])
    ref = np.array(fx["density"])
    out = y > FIX["diameter"] / 2
    contact = np.argmax(out)
    assert abs(prof[contact] / ref[contact] - 1) > 0.02
    # and it sits next to the Mark I contact value instead
    wb = FIX["profiles"]["WB-0.40"]["density"][contact]
    assert abs(prof[contact] / wb - 1) < 1e-3


@pytest.mark.parametrize("key", TENSOR_KEYS)
def test_published_tensor_profiles_coincide_with_white_bear(gpu, solute, key):
    fx = FIX["profiles"][key]
    y = np.array(fx["y"])
    ref = np.array(fx["density"])
    out = y > FIX["diameter"] / 2
    s, prof = _profile(solute, fx["eta"], deft_amd.WHITE_BEAR)
    rel = np.abs(prof[out] - ref[out]) / ref[out]
    print(key, "against Mark I: max rel", rel.max(), "mean rel", rel.mean(), flush=True)
    assert rel.max() < 4e-4 and rel.mean() < 5e-5
    s, prof_t = _profile(solute, fx["eta"], deft_amd.WHITE_BEAR_TENSOR)
    assert s["stop_reason"] == "converged", s
    rel_t = np.abs(prof_t[out] - ref[out]) / ref[out]
    print(key, "against the tensor functional: max rel", rel_t.max(), flush=True)
    assert 3e-4 < rel_t.max() < 5e-3
