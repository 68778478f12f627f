"""The measured workloads end to end on the GPU against the FP64 oracle's goldens
(tests/golden/oracle_*.json, written by tests/golden/make_oracle_golden.py on the CPU).

C2 (the headline): 128^3 solute at eta = 0.40 (excluded-centre radius 3), the bench's
    configuration -- FP32 -> FP64 ladder (switch at 1e-5 |F|, FP64 polish gradients),
    preconditioned, the reference driver's precision |p V| 1e-10 -- from the plain and
    from the network start.  Bar: |dF/F| < 1e-9.  The density is first order in the
    remaining error where the energy is second order: an energy error dF leaves
    relative density errors of order sqrt(2 dF / (N kT)) with N = 195 particles in
    the cell.  At the rule (dF <= 1.4e-7) that is 4e-5, so the density along the x
    axis is held to 1e-4 of its maximum (0.8e-5 - 3.4e-5 seen); the golden itself
    stopped at 1e-12 |F| (dF <= 1.3e-9, density errors of order 4e-6), so a run to
    relative precision 1e-12 is held to 5e-6 of it (0.9e-6 - 1.1e-6 seen).
C5: instance 0 of the seed-2024 256^3 porous batch at full size (scaled network start,
    evaluation parity with the oracle, descent); see its test for why no converged
    figure is tested.
C3: three instances of the seed-12345 64-grid batch with FP64 goldens computed from
    the unrounded potential: (96, 48, 64) at eta 0.36, (32, 64, 48) at 0.44 and
    (48, 48, 48) at 0.32, minimized in one batch with the same ladder and rule.  Bar:
    |dF/F| < 1e-9 each, and bit-identical energies and iteration counts to the same
    minimizations run one by one.
"""

import json
import math
import os

import numpy as np
import pytest

import deft_amd.workloads as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
C3_GOLDEN = (12, 6, 1)


def _gold(name):
    return json.load(open(os.path.join(GOLDEN, "oracle_%s.json" % name)))


def _ladder(dm, f, precision=None, rp=None):
    m = dm.Minimize(f)
    m.precondition(True)
    if precision is not None:
        m.set_precision(precision)
    if rp is not None:
        m.set_relative_precision(rp)
    m.set_mixed_precision(1e-5, dm.FFT_FP64)
    m.set_maxiter(1000)
    return m


@pytest.fixture(scope="module")
def c2_fluid(gpu):
    gold = _gold("c2")
    nb = 0.40 / (4 * math.pi / 3)
    f = gpu.WhiteBearFluidVeff((128, 128, 128), np.diag([12.8] * 3), mu=gpu.find_chemical_potential(nb),
                               fft_precision=gpu.FFT_FP32)
    f.Vext = W.c2_solute_vext(128, 12.8, 3.0)
    yield f, gold, nb
    f.close()


@pytest.mark.parametrize("start", ["plain", "warm"])
def test_c2_headline_meets_the_fp64_oracle(gpu, c2_fluid, start):
    import deft_amd.nn as NN
    f, gold, nb = c2_fluid
    print("c2", start, flush=True)
    if start == "warm":
        net = NN.WarmStartNet.load()
        net.warm_start(f, 0.40)
        net.close()
    else:
        f.fill_Veff(-math.log(nb))
    precision = abs(gpu.pressure(nb) * 12.8 ** 3) * 1e-10
    m = _ladder(gpu, f, precision=precision)
    m.run()
    s = m.stats()
    E = m.energy()
    rel = abs(E - gold["energy"]) / abs(gold["energy"])
    n = f.density()
    dn = float(np.max(np.abs(n[:, 0, 0] - np.array(gold["density_x_axis"]))) / gold["density_max"])
    print("c2", start, s["iter"], s["stop_reason"], "dF/F", rel, "dn", dn, m.gradient_check(), flush=True)
    assert s["stop_reason"] == "converged"
    assert rel < 1e-9
    assert dn < 1e-4
    # particle number (a whole-field checksum)
    assert abs(n.sum() / gold["density_sum"] - 1) < 1e-6
    m.close()


def test_c2_density_at_tight_precision(gpu, c2_fluid):
    f, gold, nb = c2_fluid
    f.fill_Veff(-math.log(nb))
    m = _ladder(gpu, f, rp=1e-12)
    m.run()
    n = f.density()
    dn = float(np.max(np.abs(n[:, 0, 0] - np.array(gold["density_x_axis"]))) / gold["density_max"])
    rel = abs(m.energy() - gold["energy"]) / abs(gold["energy"])
    print("c2 tight", m.stats()["iter"], "dF/F", rel, "dn", dn, flush=True)
    assert m.stats()["stop_reason"] == "converged"
    assert rel < 1e-11 and dn < 5e-6
    # N is first order in the remaining error of both runs (the golden stopped at 1e-12 |F| too)
    assert abs(n.sum() / gold["density_sum"] - 1) < 1e-7
    m.close()


def _c3_fluid(dm, inst):
    shape, cell, V, eta = inst
    nb = eta / (4 * math.pi / 3)
    f = dm.WhiteBearFluidVeff(shape, np.diag(cell), mu=dm.find_chemical_potential(nb), fft_precision=dm.FFT_FP32)
    f.Vext = V
    f.fill_Veff(-math.log(nb))
    return f, abs(dm.pressure(nb) * np.prod(cell)) * 1e-10


def test_c3_instances_batched_meet_the_fp64_oracle(gpu):
    batch = W.c3_batch(64, seed=12345)
    insts = [batch[i] for i in C3_GOLDEN]
    golds = [_gold("c3_%d" % i) for i in C3_GOLDEN]
    for inst, g in zip(insts, golds):
        assert list(inst[0]) == g["shape"]
    # one by one
    alone = []
    for inst in insts:
        f, prec = _c3_fluid(gpu, inst)
        m = _ladder(gpu, f, precision=prec)
        m.run()
        alone.append((m.stats()["iter"], m.energy()))
        m.close()
        f.close()
    # batched
    pairs = [_c3_fluid(gpu, inst) for inst in insts]
    fluids = [p[0] for p in pairs]
    mins = [_ladder(gpu, f, precision=prec) for f, prec in pairs]
    b = gpu.Batch(fluids)
    b.minimize(mins)
    for k, (m, g, a) in enumerate(zip(mins, golds, alone)):
        s = m.stats()
        E = m.energy()
        rel = abs(E - g["energy"]) / abs(g["energy"])
        n = fluids[k].density()
        print("c3", C3_GOLDEN[k], g["shape"], round(g["eta"], 3), s["iter"], s["stop_reason"], "dF/F", rel,
              "N", n.sum() / g["density_sum"] - 1, flush=True)
        assert s["stop_reason"] == "converged"
        assert rel < 1e-9
        assert abs(n.sum() / g["density_sum"] - 1) < 1e-5
        assert (s["iter"], E) == a
    b.close()
    for m in mins:
        m.close()
    for f in fluids:
        f.close()


def test_c5_instance_at_full_size(gpu):
    """C5 at its real shape: instance 0 of the seed-2024 batch (256^3, dx = 0.05, 30 hard
    spherical obstacles, eta = 0.45).  The network start inferred on the 2x coarser grid and
    upsampled has a lower FP64 energy than the bulk start; one FP64 energy + gradient at that
    start meets the FP64 oracle (|dF/F| < 1e-12, gradient 1e-10 of its scale); 30 ladder
    iterations lower the FP64 energy further.  C5 minimizations do not converge within 2000
    iterations (slow relaxation in the pores), so no converged figure is tested."""
    import deft_amd.nn as NN
    from oracle import fmt_oracle as O
    shape, cell, V, eta = W.c5_batch(1, seed=2024)[0]
    nb = W.bulk_density(eta)
    print("c5 instance 0", shape, "excluded", float(np.mean(np.isinf(V))), flush=True)
    f = gpu.WhiteBearFluidVeff(shape, np.diag(cell), mu=gpu.find_chemical_potential(nb), fft_precision=gpu.FFT_FP64)
    f.Vext = V
    f.fill_Veff(-math.log(nb))
    E_bulk = f.energy()
    net = NN.WarmStartNet.load()
    lam = net.warm_start(f, eta, factor=2)
    net.close()
    f.fft_precision = gpu.FFT_FP64
    E, g, _ = f.energy_grad_and_precond()
    print("c5 bulk start", E_bulk, "network start", E, "blend", lam, flush=True)
    assert np.isfinite(E) and E < E_bulk
    grid = O.Grid(shape, np.diag(cell))
    of = O.make_fluid(grid, eta, vext=V)
    of.x = f.Veff
    Eo, go, _ = of.energy_grad_and_precond()
    print("c5 oracle", Eo, "dF/F", abs(E - Eo) / abs(Eo), flush=True)
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go))
    del of, go, g
    m = _ladder(gpu, f, precision=abs(gpu.pressure(nb) * np.prod(cell)) * 1e-10)
    m.set_maxiter(30)
    m.run()
    f.fft_precision = gpu.FFT_FP64
    E30 = f.energy()
    print("c5 after", m.stats()["iter"], "iterations", E30, m.stats()["stop_reason"], flush=True)
    assert E30 < E - 1e-6 * abs(E)
    m.close()
    f.close()
