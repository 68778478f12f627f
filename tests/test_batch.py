"""Batched minimizations of uneven grids (deft_batch_*): every grid of a batch must
follow its own unbatched trajectory bit for bit, and the batch must run fewer
launches than the grids would one by one."""

import math

import numpy as np
import pytest

import deft_amd
import deft_amd.workloads as W

pytestmark = pytest.mark.gpu


def small_c3(count, seed, edges=(32, 48)):
    """C3 instances with test-sized edges."""
    rng = np.random.default_rng(seed)
    return [W.c3_instance(rng, edges) for _ in range(count)]


def make_fluid(inst, prec):
    shape, cell, V, eta = inst
    nb = eta / (4 * math.pi / 3)
    f = deft_amd.WhiteBearFluidVeff(shape, np.diag(cell), R=1.0, kT=1.0, mu=deft_amd.find_chemical_potential(nb),
                                    fft_precision=prec)
    f.Vext = V
    f.Veff = np.full(shape, -math.log(nb))
    return f


def make_min(f, ladder):
    m = deft_amd.Minimize(f)
    m.precondition(True)
    m.set_relative_precision(1e-10)
    if ladder:
        m.set_mixed_precision(1e-5, deft_amd.FFT_FP64)
    m.set_maxiter(300)
    return m


@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("ladder", [True, False])
def test_batch_follows_unbatched_trajectories(gpu, ladder, graphs):
    """With graphs the same captured graph is replayed every round, with each
    grid's own eager work (CG updates, reductions) between the replays; every
    replay's energies must equal the unbatched ones bit for bit."""
    insts = small_c3(5, seed=7)
    prec = deft_amd.FFT_FP32 if ladder else deft_amd.FFT_FP64
    alone = []
    for inst in insts:
        f = make_fluid(inst, prec)
        m = make_min(f, ladder)
        m.run()
        alone.append((m.stats(), m.energy(), f.Veff.copy()))
        m.close()
        f.close()
    fluids = [make_fluid(inst, prec) for inst in insts]
    mins = [make_min(f, ladder) for f in fluids]
    b = deft_amd.Batch(fluids, graphs=graphs)
    b.minimize(mins)
    st = b.stats()
    for i, (m, f) in enumerate(zip(mins, fluids)):
        s = m.stats()
        sa, ea, va = alone[i]
        assert s["iter"] == sa["iter"], (i, s, sa)
        assert s["stop_reason"] == sa["stop_reason"]
        assert s["num_energy_calcs"] == sa["num_energy_calcs"] and s["num_grad_calcs"] == sa["num_grad_calcs"]
        assert m.energy() == ea, (i, m.energy(), ea)
    total_evals = sum(a[0]["num_energy_calcs"] + a[0]["num_grad_calcs"] for a in alone)
    assert st["requests"] >= max(a[0]["num_energy_calcs"] for a in alone)
    assert st["rounds"] < st["requests"]
    assert st["rounds"] <= total_evals
    if graphs:
        # a second minimization on the same batch replays the graphs captured by the first
        for f, inst in zip(fluids, insts):
            f.Veff = np.full(inst[0], -math.log(inst[3] / (4 * math.pi / 3)))
        mins2 = [make_min(f, ladder) for f in fluids]
        b.minimize(mins2)
        for i, m in enumerate(mins2):
            assert m.energy() == alone[i][1], i
            m.close()
    b.close()
    for i, f in enumerate(fluids):
        assert np.array_equal(f.Veff, alone[i][2]), i  # the plan is usable again after the batch
    for m in mins:
        m.close()
    for f in fluids:
        f.close()


def test_batch_misuse_is_rejected(gpu):
    insts = small_c3(2, seed=3)
    f1, f2, f3 = (make_fluid(insts[0], 0), make_fluid(insts[1], 0), make_fluid(insts[0], 0))
    b = deft_amd.Batch([f1, f2])
    with pytest.raises(deft_amd.DeftError):
        deft_amd.Batch([f2])  # already in a batch
    with pytest.raises(deft_amd.DeftError):
        b.minimize([make_min(f3, True)])  # plan outside the batch
    m1 = make_min(f1, True)
    with pytest.raises(deft_amd.DeftError):
        b.minimize([m1, make_min(f1, True)])  # two minimizers on one plan
    # a plan in a batch cannot be destroyed
    assert deft_amd.lib().deft_plan_destroy(f1.handle) != 0
    b.close()
    for f in (f1, f2, f3):
        f.close()


@pytest.mark.parametrize("ladder", [True, False])
def test_batch_structure_factor_follows_unbatched(gpu, ladder):
    """The structure-factor preconditioner in batched rounds: the same fused
    passes per grid, so every grid follows its unbatched trajectory bit for bit."""
    insts = small_c3(4, seed=11)
    prec = deft_amd.FFT_FP32 if ladder else deft_amd.FFT_FP64
    alone = []
    for inst in insts:
        f = make_fluid(inst, prec)
        m = make_min(f, ladder)
        m.set_preconditioner(deft_amd.PGRAD_STRUCTURE_FACTOR)
        m.run()
        alone.append((m.stats(), m.energy()))
        m.close()
        f.close()
    fluids = [make_fluid(inst, prec) for inst in insts]
    mins = [make_min(f, ladder) for f in fluids]
    for m in mins:
        m.set_preconditioner(deft_amd.PGRAD_STRUCTURE_FACTOR)
    b = deft_amd.Batch(fluids, graphs=False)
    b.minimize(mins)
    for i, m in enumerate(mins):
        s, (sa, ea) = m.stats(), alone[i]
        assert s["iter"] == sa["iter"], (i, s, sa)
        assert s["stop_reason"] == sa["stop_reason"]
        assert s["num_energy_calcs"] == sa["num_energy_calcs"] and s["num_grad_calcs"] == sa["num_grad_calcs"]
        assert m.energy() == ea, (i, m.energy(), ea)
    b.close()
    for m in mins:
        m.close()
    for f in fluids:
        f.close()


def test_batch_graph_rounds_refuse_structure_factor(gpu):
    insts = small_c3(2, seed=5)
    fluids = [make_fluid(inst, deft_amd.FFT_FP64) for inst in insts]
    mins = [make_min(f, False) for f in fluids]
    for m in mins:
        m.set_preconditioner(deft_amd.PGRAD_STRUCTURE_FACTOR)
    b = deft_amd.Batch(fluids, graphs=True)
    with pytest.raises(deft_amd.DeftError):
        b.minimize(mins)
    b.close()
    for m in mins:
        m.close()
    for f in fluids:
        f.close()
