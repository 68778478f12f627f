"""CPU checks of the product library: it loads, exports what include/deft_amd.h declares,
and its host-side argument checks and bulk path work without a GPU."""

import math

import numpy as np
import pytest

import deft_amd
from oracle import fmt_oracle as O


def test_library_exports_every_declared_symbol():
    L = deft_amd.lib()
    declared = deft_amd.declared_symbols()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(L, name), name
    # and the Python binding covers all of them
    assert set(declared) == set(deft_amd._SIGNATURES)


def test_supported_lengths():
    """Fused per-edge kernels for ten edges; the generic line passes for any other edge
    up to 4096 whose prime factors are at most 512 (FFTW's any-N in the reference)."""
    for n in (16, 20, 21, 32, 48, 64, 96, 128, 201, 256):
        assert deft_amd.supported_length(n) and deft_amd.fused_length(n)
    for n in (1, 15, 18, 65, 100, 512, 3601, 4096):
        assert deft_amd.supported_length(n) and not deft_amd.fused_length(n)
    for n in (0, -3, 4097, 1031, 2 * 521):
        assert not deft_amd.supported_length(n) and not deft_amd.fused_length(n)


@pytest.mark.parametrize("eta", [0.1, 0.3, 0.4, 0.45])
def test_bulk_path_matches_oracle(eta):
    n = O.bulk_density(eta, 1.0)
    assert abs(deft_amd.find_chemical_potential(n) / O.chemical_potential(n, 1.0) - 1) < 1e-13
    assert abs(deft_amd.pressure(n) / O.pressure(n, 1.0) - 1) < 1e-12
    f = 1.0 * n * (math.log(n / O.nQ(1.0)) - 1) + O.homogeneous_whitebear(n, 1.0)
    assert abs(deft_amd.bulk_free_energy_density(n) / f - 1) < 1e-13


@pytest.mark.parametrize("functional,name", [(0, "wb"), (1, "wb2"), (2, "wb2slow"), (3, "wbt")])
@pytest.mark.parametrize("eta", [0.1, 0.4])
def test_bulk_state_of_every_functional(eta, functional, name):
    n = O.bulk_density(eta, 1.0)
    f, mu, p = deft_amd.bulk_state(n, 1.0, 1.0, functional)
    assert abs(mu / O.chemical_potential(n, 1.0, functional=name) - 1) < 1e-13
    assert abs(p / O.pressure(n, 1.0, functional=name) - 1) < 1e-12
    fo = n * (math.log(n / O.nQ(1.0)) - 1) + O.homogeneous_whitebear(n, 1.0, functional=name)
    assert abs(f / fo - 1) < 1e-13
    assert deft_amd.find_chemical_potential(n, functional=functional) == mu
    with pytest.raises(deft_amd.DeftError):
        deft_amd.bulk_state(n, 1.0, 1.0, 4)


def test_bad_arguments_raise_without_device():
    with pytest.raises(deft_amd.DeftError):
        deft_amd.find_chemical_potential(-1.0)
    # without a GPU plan creation fails loudly; with one, an edge with a prime factor
    # above 512 is rejected
    with pytest.raises(deft_amd.DeftError):
        deft_amd.WhiteBearFluidVeff((1031, 16, 16), np.diag([1.0, 1.0, 1.0]))


def test_product_path_has_no_oracle_import():
    import ast
    import os
    pkg = os.path.dirname(deft_amd.__file__)
    for root, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith(".py"):
                tree = ast.parse(open(os.path.join(root, fn)).read())
                for node in ast.walk(tree):
                    if isinstance(node, (ast.Import, ast.ImportFrom)):
                        names = [a.name for a in node.names] + [getattr(node, "module", "") or ""]
                        assert not any(n.startswith("oracle") for n in names), fn
