"""Benchmark: converged White-Bear minimizations per second at 128^3 on one MI355X,
from the warm-start network's prediction and from the plain start.

Workload (a "step" = one complete minimization, the warm start's network
inference included): hard-sphere fluid, R = 1, kT = 1, bulk packing fraction
0.40, around a spherical hard solute (Vext = inf within radius 3 of the origin,
periodic) in a 12.8^3 cell on a 128^3 grid (dx = 0.1); mu from the bulk
equation of state; start Veff = -kT (ln n_b + u) with u from the committed
network (deft_amd/models/unet_warmstart.npz), or Veff = -kT ln n_b (plain);
Polak-Ribiere CG with deft's quadratic line search, preconditioned with the
structure factor of the bulk fluid (value) or with the reference's grad/n
(reported beside it); FP32 FFTs /
FP64 pointwise and reductions until the error estimate is below SWITCH_RP |F|,
then FP64 throughout to the reference driver's absolute precision |p_bulk V| 1e-10
(about 1e-10 |F|), with the quadratic step's
trial point evaluated for energy and gradient at once.  Every step is checked
against the FP64 oracle's converged free energy (|dF/F| < 1e-9).

Prints one JSON line (see the "roofline" and "cpu_baseline" objects).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--cpu-sample 15]
"""

import argparse
import json
import math
import os
import platform
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import deft_amd as D  # noqa: E402
import deft_amd.nn as NN  # noqa: E402
import deft_amd.workloads as W  # noqa: E402

METRIC = "converged FMT minimizations/s at 128^3, NN warm-start vs plain, |dF/F|<1e-9"
N, L, ETA, R, KT, SOLUTE = 128, 12.8, 0.40, 1.0, 1.0, 3.0
SWITCH_RP = 1e-5
POLISH_GRAD = D.FFT_FP64
# convergence rule of the reference driver (papers/contact/figs/inner-sphere.cpp:201-203):
# absolute precision |omega_bulk V| 1e-10 with omega_bulk = -p the bulk grand free energy density
PRECISION_FACTOR = 1e-10
HBM_PEAK_GBS = 8000.0


def gpu_run(steps, warmup):
    nb = ETA / (4 * math.pi / 3 * R ** 3)
    mu = D.find_chemical_potential(nb, R, KT)
    f = D.WhiteBearFluidVeff((N, N, N), np.diag([L] * 3), R=R, kT=KT, mu=mu, fft_precision=D.FFT_FP32)
    f.Vext = W.c2_solute_vext(N, L, SOLUTE)
    tuning = f.autotune()
    eng = f.engine()
    print("autotuned pass variants:", tuning, "engine:", eng, file=sys.stderr, flush=True)
    out_tuning = tuning
    net = NN.WarmStartNet.load()
    v0 = -KT * math.log(nb)
    gold = json.load(open(os.path.join(HERE, "tests", "golden", "oracle_c2.json")))
    precision = abs(D.pressure(nb, R, KT) * L ** 3) * PRECISION_FACTOR

    def one(warm, sk):
        f.fft_precision = D.FFT_FP32
        t0 = time.perf_counter()
        if warm:
            net.warm_start(f, ETA)
        else:
            f.fill_Veff(v0)
        t1 = time.perf_counter()
        m = D.Minimize(f)
        m.precondition(True)
        if sk:
            m.set_preconditioner(D.PGRAD_STRUCTURE_FACTOR)
        m.set_precision(precision)
        m.set_mixed_precision(SWITCH_RP, POLISH_GRAD)
        m.run()
        t2 = time.perf_counter()
        s = m.stats()
        s["fp32_iterations"], s["mixed_iterations"], _ = m.phase_iterations()
        s["seconds"] = t2 - t0
        s["start_seconds"] = t1 - t0
        s["energy"] = m.energy()
        m.close()
        return s

    out = {}
    for sk in (True, False):
        for warm in (True, False):
            for _ in range(warmup):
                one(warm, sk)
            results = []
            t0 = time.perf_counter()
            for _ in range(steps):
                results.append(one(warm, sk))
            elapsed = time.perf_counter() - t0
            n = f.density()
            dn = float(np.max(np.abs(n[:, 0, 0] - np.array(gold["density_x_axis"]))) / gold["density_max"])
            errs = [abs(r["energy"] - gold["energy"]) / abs(gold["energy"]) for r in results]
            out[("warm" if warm else "plain") + ("" if sk else "_ref")] = dict(results=results, elapsed=elapsed,
                                                                               err=max(errs), dn=dn)
    # the value's steps again with HIP events around every launch (per-kernel times
    # for the roofline; kept out of the throughput timing above)
    f.set_profiling(True)
    f.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        one(True, True)
    out["profiled_elapsed"] = time.perf_counter() - t0
    out["tuning"] = out_tuning
    out["engine"] = {"chosen": eng[0], "fused_eg_ms_fp32_plus_fp64": eng[1], "generic_eg_ms_fp32_plus_fp64": eng[2]}
    out["prof"] = f.profile()
    f.set_profiling(False)
    return out, gold, precision


def c3_run(count=64, seed=12345, reps=2, precond=D.PGRAD_STRUCTURE_FACTOR, sequential=True):
    """64 uneven grids (edges per axis from {32, 48, 64, 96}, dx = 0.1, 8 Gaussians and a
    planar hard wall, eta ~ U(0.1, 0.45)): one batch against the same minimizations run
    one after another; both from the bulk start, same ladder and precision rule as above,
    preconditioned with `precond` (the structure factor, or deft's grad/n)."""
    insts = W.c3_batch(count, seed=seed)
    fluids = []
    for shape, cell, V, eta in insts:
        nb = eta / (4 * math.pi / 3 * R ** 3)
        f = D.WhiteBearFluidVeff(shape, np.diag(cell), R=R, kT=KT, mu=D.find_chemical_potential(nb, R, KT),
                                 fft_precision=D.FFT_FP32)
        f.Vext = V
        fluids.append((f, -KT * math.log(nb), abs(D.pressure(nb, R, KT) * np.prod(cell)) * PRECISION_FACTOR))

    def minimizers():
        ms = []
        for f, v0, prec in fluids:
            f.fill_Veff(v0)
            m = D.Minimize(f)
            m.precondition(True)
            if precond != D.PGRAD_NEW:
                m.set_preconditioner(precond)
            m.set_precision(prec)
            m.set_mixed_precision(SWITCH_RP, POLISH_GRAD)
            ms.append(m)
        return ms

    def outcome(ms):
        r = [(m.stats()["iter"], m.energy(), m.stats()["stop_reason"]) for m in ms]
        for m in ms:
            m.close()
        return r

    best_seq = best_bat = seq = None
    for _ in range(reps if sequential else 0):
        ms = minimizers()
        t0 = time.perf_counter()
        for m in ms:
            m.run()
        dt = time.perf_counter() - t0
        seq = outcome(ms)
        best_seq = dt if best_seq is None else min(best_seq, dt)
    batch = D.Batch([f for f, _, _ in fluids])
    for _ in range(reps):
        ms = minimizers()
        t0 = time.perf_counter()
        batch.minimize(ms)
        dt = time.perf_counter() - t0
        bat = outcome(ms)
        best_bat = dt if best_bat is None else min(best_bat, dt)
    st = batch.stats()
    batch.close()
    for f, _, _ in fluids:
        f.close()
    res = {"workload": f"{count} uneven grids, edges per axis from {{32, 48, 64, 96}}, seed {seed}",
           "preconditioner": "grad/n" if precond == D.PGRAD_NEW else "structure factor",
           "points": int(sum(np.prod(s) for s, _, _, _ in insts)),
           "batched_minimizations_per_s": count / best_bat,
           "rounds_per_batch": st["rounds"] // reps, "launches_per_batch": st["launches"] // reps,
           "iterations": [b[0] for b in bat], "total_iterations": int(sum(b[0] for b in bat)),
           "all_converged": all(b[2] == "converged" for b in bat)}
    if sequential:
        res.update({"sequential_minimizations_per_s": count / best_seq, "batched_over_sequential": best_seq / best_bat,
                    "identical_to_sequential": all(a[:2] == b[:2] for a, b in zip(seq, bat))})
    return res


def c4_run(count=512, chunk=64, seed=4242, M=64, L=6.4):
    """The learned warm start on held-out potentials (C4): `count` potentials from the
    training generator (deft_amd.nn.training_vext: hard spheres, slabs and Gaussian
    bumps, or the C3 Gaussians and wall) at 64^3, dx = 0.1, eta ~ U(0.1, 0.45), drawn
    from a seed the training never used; each minimized from the plain (bulk) start and
    from the network's start, in batches of `chunk`, with the headline's ladder and
    precision rule and the structure-factor preconditioner.
    Times count the start (bulk fill, or one network inference over the chunk) and the
    minimization."""
    rng = np.random.default_rng(seed)
    cases = [(NN.training_vext(M, L, rng), float(rng.uniform(0.1, 0.45))) for _ in range(count)]
    net = NN.WarmStartNet.load()
    fluids = [D.WhiteBearFluidVeff((M, M, M), np.diag([L] * 3), R=R, kT=KT, mu=0.0, fft_precision=D.FFT_FP32)
              for _ in range(chunk)]
    out = {k: {"iterations": [], "fp32_iterations": [], "seconds": 0.0, "converged": 0, "energies": []}
           for k in ("plain", "warm")}
    for c0 in range(0, count, chunk):
        part = cases[c0:c0 + chunk]
        fl = fluids[:len(part)]
        precs, nbs = [], []
        for f, (V, eta) in zip(fl, part):
            nb = eta / (4 * math.pi / 3 * R ** 3)
            f.mu = D.find_chemical_potential(nb, R, KT)
            f.Vext = V
            nbs.append(nb)
            precs.append(abs(D.pressure(nb, R, KT) * L ** 3) * PRECISION_FACTOR)
        for start in ("plain", "warm"):
            t0 = time.perf_counter()
            for f in fl:
                f.fft_precision = D.FFT_FP32
            if start == "warm":
                # one inference over the whole chunk (deft_plans_warm_start)
                NN.warm_start_batch(net, fl, [eta for _, eta in part])
            else:
                for f, nb in zip(fl, nbs):
                    f.fill_Veff(-KT * math.log(nb))
            batch = D.Batch(fl)  # after the warm starts: their safety check evaluates outside a batch
            ms = []
            for f, prec in zip(fl, precs):
                m = D.Minimize(f)
                m.precondition(True)
                m.set_preconditioner(D.PGRAD_STRUCTURE_FACTOR)
                m.set_precision(prec)
                m.set_maxiter(1000)
                m.set_mixed_precision(SWITCH_RP, POLISH_GRAD)
                ms.append(m)
            batch.minimize(ms)
            out[start]["seconds"] += time.perf_counter() - t0
            for m in ms:
                st = m.stats()
                out[start]["iterations"].append(st["iter"])
                out[start]["fp32_iterations"].append(m.phase_iterations()[0])
                out[start]["converged"] += st["stop_reason"] == "converged"
                out[start]["energies"].append(m.energy())
                m.close()
            batch.close()
    net.close()
    for f in fluids:
        f.close()
    res = {"workload": f"{count} held-out potentials of the training generator at {M}^3, dx = {L / M:g}, "
                       f"eta ~ U(0.1, 0.45), seed {seed}; batches of {chunk}; structure-factor preconditioner"}
    for k in ("plain", "warm"):
        it = np.array(out[k]["iterations"])
        res[k] = {"mean_iterations": float(it.mean()), "median_iterations": float(np.median(it)),
                  "mean_fp32_iterations": float(np.mean(out[k]["fp32_iterations"])),
                  "minimizations_per_s": count / out[k]["seconds"], "converged": out[k]["converged"]}
    ep, ew = np.array(out["plain"]["energies"]), np.array(out["warm"]["energies"])
    res["iteration_ratio_plain_over_warm"] = res["plain"]["mean_iterations"] / res["warm"]["mean_iterations"]
    res["time_ratio_plain_over_warm"] = res["warm"]["minimizations_per_s"] / res["plain"]["minimizations_per_s"]
    res["max_dF_over_F_between_starts"] = float(np.max(np.abs(ew - ep) / np.abs(ep)))
    return res


def c5_run(iters=40, seed=2024):
    """C5 at its real shape: instance 0 of the seed-2024 porous batch (256^3, dx = 0.05,
    30 hard spherical obstacles of excluded radius U(1, 2.5), eta = 0.45), started from the
    network's prediction on the 2x coarser grid, upsampled, then the headline's ladder for
    `iters` CG iterations.  C5 minimizations do not converge within 2000 iterations with
    either preconditioner (slow relaxation in the pores), so C5 is reported as iteration
    throughput and the FP64 energy it reaches, not as minimizations per second."""
    shape, cell, V, eta = W.c5_batch(1, seed=seed)[0]
    nb = W.bulk_density(eta)
    f = D.WhiteBearFluidVeff(shape, np.diag(cell), R=R, kT=KT, mu=D.find_chemical_potential(nb, R, KT),
                             fft_precision=D.FFT_FP32)
    f.Vext = V
    f.autotune()
    net = NN.WarmStartNet.load()
    precision = abs(D.pressure(nb, R, KT) * np.prod(cell)) * PRECISION_FACTOR

    def one():
        f.fft_precision = D.FFT_FP32
        t0 = time.perf_counter()
        net.warm_start(f, eta, factor=2)
        t1 = time.perf_counter()
        m = D.Minimize(f)
        m.precondition(True)
        m.set_precision(precision)
        m.set_maxiter(iters)
        m.set_mixed_precision(SWITCH_RP, POLISH_GRAD)
        m.run()
        t2 = time.perf_counter()
        st = m.stats()
        m.close()
        return t1 - t0, t2 - t1, st

    one()
    tn, tm, st = one()
    f.fft_precision = D.FFT_FP64
    E_end = f.energy()
    net.warm_start(f, eta, factor=2)
    f.fft_precision = D.FFT_FP64
    E_start = f.energy()
    net.close()
    f.close()
    return {"workload": f"C5 instance 0 of seed {seed}: 256^3, dx = 0.05, 30 hard obstacles, eta = 0.45, "
                        f"network start on the 2x coarser grid, {iters} ladder iterations (grad/n)",
            "excluded_fraction": float(np.mean(np.isinf(V))), "iterations": st["iter"],
            "ms_per_iteration": tm / st["iter"] * 1e3,
            "iterations_per_s": st["iter"] / tm, "network_ms": tn * 1e3,
            "energy_calcs": st["num_energy_calcs"], "grad_calcs": st["num_grad_calcs"],
            "fp64_energy_at_start": E_start, "fp64_energy_after": E_end, "stop_reason": st["stop_reason"],
            "converged": False}


def event_roofline(f, run):
    """Runs `run()` with the plan's HIP-event profiler on and returns, per kernel
    family with time, its launches, ms, algorithmic GB/s and fraction of the
    8 TB/s peak (the families and byte counts of ProfScope, plan.hpp)."""
    f.profile_reset()
    f.set_profiling(True)
    run()
    f.set_profiling(False)
    out = {}
    for k, (n, ms, by) in f.profile().items():
        if n > 0 and ms > 0:
            gbs = by / (ms * 1e-3) / 1e9
            out[k] = {"launches": n, "ms": round(ms, 3), "GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 3)}
    return dict(sorted(out.items(), key=lambda kv: -kv[1]["ms"]))


def sw_run(kT=5.0, ff=0.3, lam=1.3):
    """The square-well liquid around a test particle at its published shape
    (papers/square-well-fluid/figs/new-radial-sw.cpp: 18^3 cell, dx = 0.1 ->
    180^3, R = 1, epsilon = 1, lambda = 1.3, a 400 kT core, relative precision 0,
    at most 500 iterations, preconditioned, FP64): one minimization, timed, and
    its deviation from the reference's published profile of that case."""
    R, eps, edge, dx = 1.0, 1.0, 18.0, 0.1
    sigma = 2 * R
    Ng = int(math.ceil(edge / dx / 2)) * 2
    idx = np.arange(Ng)
    c = np.where(idx > Ng // 2, idx - Ng, idx) * edge / Ng
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    nb = ff / (math.pi * sigma ** 3 / 6)
    vmax = 400 * kT
    vext = np.where(r <= sigma, vmax, np.where((r > sigma) & (r < sigma * lam), -eps, 0.0))
    veff = -kT * math.log(nb) + np.where(r <= sigma, vmax, 0.0)
    mu = D.HomogeneousSW_liquid(R=R, epsilon=eps, lambda_=lam, kT=kT, n=nb).d_by_dn()
    f = D.SWLiquidVeff((Ng, Ng, Ng), np.diag([edge] * 3), R=R, epsilon=eps, lambda_=lam, kT=kT, mu=mu,
                       fft_precision=D.FFT_FP64)
    f.autotune()
    f.Vext = vext

    def one():
        f.Veff = veff
        m = D.Minimize(f)
        m.set_relative_precision(0)
        m.set_maxiter(500)
        m.precondition(True)
        t0 = time.perf_counter()
        m.run()
        dt = time.perf_counter() - t0
        st = m.stats()
        m.close()
        return dt, st

    one()
    dt, st = one()
    prof = event_roofline(f, one)
    eta = f.density()[0, 0, :Ng // 2] * math.pi * sigma ** 3 / 6
    gold = json.load(open(os.path.join(HERE, "tests", "golden", "sw_radial.json")))["cases"]
    case = [g for g in gold if g["kT"] == kT and g["ff"] == ff][0]
    pub = np.array(case["eta"])
    out = np.array(case["r_over_sigma"]) > 1.0
    f.close()
    return {"workload": f"square-well liquid around a test particle, 180^3, 18^3 cell, kT = {kT:g}, "
                        f"ff = {ff:g}, lambda = {lam:g}, the reference driver's protocol (FP64, grad/n)",
            "minimizations_per_s": 1.0 / dt, "seconds": dt, "iterations": st["iter"],
            "energy_calcs": st["num_energy_calcs"], "grad_calcs": st["num_grad_calcs"],
            "ms_per_iteration": dt / st["iter"] * 1e3,
            "max_rel_dev_from_published": float(np.max(np.abs(eta[out] - pub[out]) / pub[out])),
            "kernel_families": prof}


def sfmt_run(iters=30):
    """Soft FMT (sigma = epsilon = 1, Xi = 0.1, kT = 1) around the headline's radius-3
    hard solute at 128^3, 12.8^3 cell, bulk packing 0.40 of the sphere of diameter
    alpha: `iters` CG iterations of the FP32 -> FP64 ladder with grad/n, timed."""
    L, M, kT, xi = 12.8, 128, 1.0, 0.1
    hf = D.HomogeneousSFMTFluid(sigma=1.0, epsilon=1.0, Xi=xi, kT=kT)
    f = D.SFMTFluidVeff((M, M, M), np.diag([L] * 3), sigma=1.0, epsilon=1.0, Xi=xi, kT=kT,
                        fft_precision=D.FFT_FP32)
    alpha = f.alpha
    nb = 0.40 / (math.pi * alpha ** 3 / 6)
    hf.n = nb
    f.mu = hf.d_by_dn()
    f.Vext = W.sphere_vext((M, M, M), (L, L, L), (0.0, 0.0, 0.0), 3.0)
    f.autotune()

    def one():
        f.fft_precision = D.FFT_FP32
        f.fill_Veff(-kT * math.log(nb))
        m = D.Minimize(f)
        m.precondition(True)
        m.set_relative_precision(1e-12)
        m.set_maxiter(iters)
        m.set_mixed_precision(SWITCH_RP, POLISH_GRAD)
        t0 = time.perf_counter()
        m.run()
        dt = time.perf_counter() - t0
        st = m.stats()
        m.close()
        return dt, st

    one()
    dt, st = one()
    prof = event_roofline(f, one)
    f.close()
    return {"workload": f"soft FMT (sigma = epsilon = kT = 1, Xi = {xi:g}) around a radius-3 hard solute, "
                        f"{M}^3, {L:g}^3 cell, eta = 0.40, {iters} ladder iterations (grad/n)",
            "iterations": st["iter"], "ms_per_iteration": dt / st["iter"] * 1e3,
            "iterations_per_s": st["iter"] / dt, "energy_calcs": st["num_energy_calcs"],
            "grad_calcs": st["num_grad_calcs"], "kernel_families": prof}


def rocprof_instance(family, tuning):
    """rocprof name of the kernel instance behind a profiler family at the autotuned
    variants (`tuning` as returned by autotune()); None when there is no single one."""
    base, _, prec = family.rpartition("_")
    T = {"f32": "float", "f64": "double"}.get(prec)
    if not T:
        return None
    t = tuning[0 if prec == "f32" else 1]
    zb, ty, tx = t["zblock"], t["ytile"], t["xtile"]
    names = {"zfwd_density": f"deft::k_zfwd_density<{T}, {N}, {zb}>",
             "ypass1_fwd": f"deft::k_ypass<{T}, {N}, false, 1, {ty}>",
             "ypass1_inv": f"deft::k_ypass<{T}, {N}, true, 1, {ty}>",
             "ypass5_inv": f"deft::k_ypass<{T}, {N}, true, 5, {ty}>",
             "ypass5_fwd": f"deft::k_ypass<{T}, {N}, false, 5, {ty}>",
             "xconv_fwd": f"deft::k_xconv_fwd<{T}, {N}, {tx}, 5>", "xconv_bwd": f"deft::k_xconv_bwd<{T}, {N}, {tx}, 5>",
             "zphi_energy": f"deft::k_zphi<{T}, {N}, false, {zb}, 0>",
             "zphi_grad": f"deft::k_zphi<{T}, {N}, true, {zb}, 0>",
             "zgrad": f"deft::k_zgrad<{T}, {N}, {zb}>"}
    # the one-per-wave variants of the Phi and forward x passes
    if t.get("zphi"):
        names["zphi_energy"] = f"deft::k_zphi_wave<{T}, {N}, false>"
        names["zphi_grad"] = f"deft::k_zphi_wave<{T}, {N}, true>"
    if t.get("xfwd"):
        names["xconv_fwd"] = f"deft::k_xconv_fwd_wave<{T}, {N}>"
    return names.get(base)


def committed_traffic(family, tuning):
    """HBM bytes per dispatch of that kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (profiles/c2_128_solute_pmc.json), or None."""
    path = os.path.join(HERE, "profiles", "c2_128_solute_pmc.json")
    inst = rocprof_instance(family, tuning)
    if not inst or not os.path.exists(path):
        return None
    t = json.load(open(path))["per_dispatch_hbm_bytes"].get(inst)
    return None if t is None else t["fetch_bytes"] + t["write_bytes"]


def committed_busy_fraction():
    """GPU busy fraction of the timed warm-start steps in the committed rocprofv3 kernel trace
    summary of this workload (profiles/c2_128_solute_pmc.json), or None."""
    path = os.path.join(HERE, "profiles", "c2_128_solute_pmc.json")
    if not os.path.exists(path):
        return None
    return json.load(open(path)).get("warm_step_busy_fraction")


def whole_path_bytes(prof, steps):
    """Algorithmic HBM bytes of one minimization by the whole-path rule: every 3-D FFT
    reads its input and writes its output once (B_fft = 4N + 8Nk FP32, 8N + 16Nk FP64),
    one energy B_E = 6 B_fft + 12N (+ 24N for the line-search move it follows), one
    energy + gradient B_EG = 12 B_fft + 20N, and 40N of CG vectors per iteration; the
    evaluations are counted from the profiled steps (energy-only = Phi energy passes,
    energy + gradient = Phi gradient passes, iterations = CG updates), plus per
    structure-factor preconditioned gradient one forward and one inverse transform,
    the spectral scale (2 x 8Nk FP32 / 2 x 16Nk FP64 + 8Nk) and the last pass (48N)."""
    Nr = float(N) ** 3
    Nk = float(N) * N * (N // 2 + 1)
    total = 0.0
    for prec, bfft in (("f32", 4 * Nr + 8 * Nk), ("f64", 8 * Nr + 16 * Nk)):
        nE = prof.get("zphi_energy_" + prec, (0,))[0]
        nEG = prof.get("zphi_grad_" + prec, (0,))[0]
        total += nE * (6 * bfft + 12 * Nr + 24 * Nr) + nEG * (12 * bfft + 20 * Nr)
        nSK = prof.get("sk_finish_" + prec, (0,))[0]
        total += nSK * (2 * bfft + (bfft - (4 if prec == "f32" else 8) * Nr) + 8 * Nk + 48 * Nr)
    total += prof.get("cg_update", (0,))[0] * 40 * Nr
    return total / steps


def _cpu_info():
    model = platform.processor()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        fftw = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True).stdout
        fftw = "libfftw3" in fftw
    except OSError:
        fftw = False
    return model, os.cpu_count(), fftw


def cpu_baseline(sample_seconds):
    """deft's FP64 CPU path restated in C (oracle/fmt_cpu.c: the new API's cached-energy
    dataflow, the project's own FFT), one single-threaded minimization per process as in the
    reference, one process per host core of this GPU's CPU share, same workload, start and
    convergence rule as the GPU's plain run.  Each process runs the minimization for a bounded
    sample and reports its per-evaluation times; a minimization costs the full run's own
    evaluation counts (tests/golden/cpu_c2_rule.json: the same program run to convergence under
    the same rule) at those times.  Throughput = sum over processes of 1 / (time per minimization)."""
    exe = os.path.join(HERE, "oracle", "build", "fmt_cpu")
    if not os.path.exists(exe):
        subprocess.run(["make", "-C", os.path.join(HERE, "oracle"), "build/fmt_cpu"], check=True,
                       capture_output=True)
    full = json.load(open(os.path.join(HERE, "tests", "golden", "cpu_c2_rule.json")))
    cmd = [exe, "solute", str(N), str(L), str(ETA), str(SOLUTE), repr(full["precision"]), "0"]

    def per_min(r):
        te = r["seconds_energy"] / max(r["energy_evals"], 1)
        tg = r["seconds_grad"] / max(r["grad_evals"], 1)
        evals = r["seconds_energy"] + r["seconds_grad"]
        overhead = r["seconds"] / evals if evals > 0 else 1.0  # CG vector work between evaluations
        return (te * full["energy_evals"] + tg * full["grad_evals"]) * overhead

    alone = json.loads(subprocess.run(cmd + [str(sample_seconds / 3)], capture_output=True, text=True,
                                      check=True).stdout)
    cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    procs = [subprocess.Popen(cmd + [str(sample_seconds)], stdout=subprocess.PIPE, text=True) for _ in range(cores)]
    runs = [json.loads(p.communicate()[0]) for p in procs]
    if any(p.returncode for p in procs):
        raise RuntimeError("CPU baseline process failed")
    rate = sum(1.0 / per_min(r) for r in runs)
    model, nproc, fftw = _cpu_info()
    t1 = per_min(alone)
    return {"value": rate, "unit": "minimizations/s", "cores": cores, "kind": "port",
            "sample": f"{cores} concurrent single-threaded processes, each running the 128^3 minimization for "
                      f"{sample_seconds:.0f} s ({sum(r['energy_evals'] + r['grad_evals'] for r in runs)} evaluations "
                      f"in all), scaled by the full run's {full['energy_evals']} energy and {full['grad_evals']} "
                      f"energy+gradient evaluations ({full['iterations']} iterations) under the same rule",
            "cpu_model": model, "nproc": nproc, "processes": cores,
            "fft": "the project's own FP64 Stockham FFT" + ("" if not fftw else " (libfftw3 present but not used)")
                   + ("; FFTW is not installed" if not fftw else ""),
            "program": "oracle/fmt_cpu.c (new-API Minimize, FP64, one thread per minimization)",
            "convergence_rule": "|p_bulk V| * 1e-10 absolute, plain start",
            "iterations": full["iterations"], "dF_over_F_vs_fp64_oracle": full["dF_over_F_vs_oracle_c2"],
            "single_core_seconds_per_minimization": t1,
            "all_cores_seconds_per_minimization": cores / rate}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-sample", type=float, default=15.0)
    ap.add_argument("--no-c3", action="store_true", help="skip the 64-grid batch measurement")
    ap.add_argument("--no-c4", action="store_true", help="skip the 512 held-out warm-start measurement")
    ap.add_argument("--no-c5", action="store_true", help="skip the 256^3 porous iteration measurement")
    ap.add_argument("--no-f4", action="store_true", help="skip the soft-FMT and square-well measurements")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 512 MiB device-to-device copy kernel (calibrates rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE against a known byte count)")
    ap.add_argument("--c3-only", action="store_true",
                    help="only run the 64-grid batch once (structure factor, no sequential runs), for profiling")
    ap.add_argument("--f4-only", action="store_true",
                    help="only run the square-well and soft-FMT legs and print their results (for profiling)")
    args = ap.parse_args()
    if args.c3_only:
        print(json.dumps({"c3_batch": c3_run(reps=1, sequential=False)}), flush=True)
        return
    if args.f4_only:
        print(json.dumps({"sw_radial_180": sw_run()}), flush=True)
        print(json.dumps({"sfmt_solute_128": sfmt_run()}), flush=True)
        return
    if args.copy_probe:
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": a.numel() * 8, "copy_probe_bytes_written": b.numel() * 8,
                          "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the minimizations of this workload are independent "
              f"and shard by instance); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    runs, gold, precision = gpu_run(args.steps, args.warmup)
    # attainable HBM bandwidth on this device (2 GiB streaming copy, larger than the
    # 256 MB last-level cache), next to the spec peak
    copy_gbps = D.copy_bandwidth(1 << 31, 10)
    w, pl = runs["warm"], runs["plain"]
    wr, plr = runs["warm_ref"], runs["plain_ref"]
    value = args.steps / w["elapsed"]
    plain_value = args.steps / pl["elapsed"]
    ref_value, ref_plain_value = args.steps / wr["elapsed"], args.steps / plr["elapsed"]
    prof = runs["prof"]
    # dominant kernel: the kernel instance (one profiler entry: pass, direction and FFT
    # precision at the autotuned variant) with the most event time over the profiled steps
    kname, (kl, kms, kby) = max(((k, v) for k, v in prof.items() if v[0] > 0), key=lambda kv: kv[1][1])
    achieved = kby / (kms * 1e-3) / 1e9
    total_ms = sum(v[1] for v in prof.values())
    total_by = sum(v[2] for v in prof.values())
    path_bytes = whole_path_bytes(prof, args.steps)
    path_gbps = path_bytes / (w["elapsed"] / args.steps) / 1e9
    c3 = None if args.no_c3 else c3_run()
    if c3 is not None:
        c3["grad_n_preconditioner"] = c3_run(precond=D.PGRAD_NEW, sequential=False)
    c4 = None if args.no_c4 else c4_run()
    c5 = None if args.no_c5 else c5_run()
    sw = None if args.no_f4 else sw_run()
    sfmt = None if args.no_f4 else sfmt_run()
    cpu = cpu_baseline(args.cpu_sample) if args.cpu_sample > 0 else None
    def block(run, start):
        r = run["results"][-1]
        d = {"minimizations_per_s": args.steps / run["elapsed"], "ms_per_minimization": run["elapsed"] / args.steps * 1e3,
             "iterations": r["iter"], "fp32_iterations": r["fp32_iterations"], "energy_calcs": r["num_energy_calcs"],
             "grad_calcs": r["num_grad_calcs"], "dF_over_F_vs_fp64_oracle": run["err"],
             "max_dn_over_nmax_x_axis": run["dn"]}
        if start == "warm":
            d["network_ms"] = r["start_seconds"] * 1e3
        return d
    line = {
        "metric": METRIC, "value": value, "unit": "minimizations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": w["elapsed"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 FFT + fp64 pointwise/reductions, fp64 polish; bf16 MFMA network (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "hard-sphere fluid eta=0.40 around a radius-3 hard solute, 128^3 grid, "
                               "12.8^3 cell, network warm start (value) vs plain start",
                   "preconditioner": "structure factor of the bulk fluid (DEFT_PGRAD_STRUCTURE_FACTOR); "
                                     "'reference_preconditioner' repeats both starts with deft's grad/n",
                   "grid": [N, N, N], "cell": L, "eta": ETA, "R": R,
                   "switch_relative_precision": SWITCH_RP, "precision": precision,
                   "precision_rule": "|p_bulk V| * 1e-10 (papers/contact/figs/inner-sphere.cpp:201-203)",
                   "polish_gradients": "fp64" if POLISH_GRAD == D.FFT_FP64 else "fp32",
                   "warm": block(w, "warm"),
                   "plain": block(pl, "plain"),
                   "warm_over_plain": value / plain_value,
                   "reference_preconditioner": {"warm": block(wr, "warm"), "plain": block(plr, "plain"),
                                                "warm_over_plain": ref_value / ref_plain_value},
                   "structure_factor_over_reference_preconditioner": {"warm": value / ref_value,
                                                                      "plain": plain_value / ref_plain_value},
                   "autotuned_engine": runs["engine"],
                   "c3_batch": c3,
                   "c4_warm_start_heldout": c4,
                   "c5_porous_256": c5, "sw_radial_180": sw, "sfmt_solute_128": sfmt,
                   "oracle_iterations": gold["iterations"]},
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": committed_traffic(kname, runs["tuning"]),
                     "traffic_source": "profiles/c2_128_solute_pmc.json (rocprofv3 FETCH_SIZE+WRITE_SIZE, "
                                       "copy-calibrated, bytes per dispatch)",
                     "avg_launch_us": kms / kl * 1e3, "bytes_per_launch": kby / kl,
                     # the next kernel instances by event time, with their fractions of the
                     # peak (the dominant one can change places with a close second)
                     "top_kernels": [{"kernel": k, "ms": round(v[1], 3), "avg_launch_us": round(v[1] / v[0] * 1e3, 2),
                                      "frac": round(v[2] / (v[1] * 1e-3) / 1e9 / HBM_PEAK_GBS, 3)}
                                     for k, v in sorted(((k, v) for k, v in prof.items() if v[0] > 0),
                                                        key=lambda kv: -kv[1][1])[:5]],
                     "all_kernels_GBps": total_by / (total_ms * 1e-3) / 1e9,
                     # GPU busy fraction of the warm-start steps from the committed rocprofv3 kernel
                     # trace of this workload (kernel time over the span of each step; a lower
                     # bound, the profiler adds gaps of its own)
                     "kernel_time_fraction_of_wall": committed_busy_fraction(),
                     # the plan's HIP-event kernel time over the minimizations' own wall time without
                     # events: events lengthen every kernel a little, so this overstates the fraction
                     "kernel_event_ms_over_wall": total_ms / args.steps / (1e3 * np.mean(
                         [r["seconds"] - r["start_seconds"] for r in w["results"]])),
                     "kernel_ms_per_step": total_ms / args.steps,
                     "streaming_copy_GBps": copy_gbps, "frac_of_streaming_copy": achieved / copy_gbps,
                     "whole_path": {"bytes_per_minimization": path_bytes, "achieved_GBps": path_gbps,
                                    "frac": path_gbps / HBM_PEAK_GBS,
                                    "rule": "sum over the counted evaluations of B_E = 6 B_fft + 36N (energy + "
                                            "line-search move) and B_EG = 12 B_fft + 20N, plus 40N per CG "
                                            "iteration; B_fft = 4N + 8Nk (FP32) / 8N + 16Nk (FP64); over the "
                                            "warm-start step's wall time (network inference included)"},
                     "measured": f"HIP events around every launch over {args.steps} further warm-start steps "
                                 f"({runs['profiled_elapsed'] / args.steps * 1e3:.2f} ms per step with events)"},
        "cpu_baseline": cpu,
        "gpu_over_cpu_baseline": None if not cpu else {
            "warm": value / cpu["value"], "plain": plain_value / cpu["value"],
            "plain_reference_preconditioner": ref_plain_value / cpu["value"],
            "note": "the CPU baseline runs deft's grad/n preconditioner; plain_reference_preconditioner is "
                    "the like-for-like algorithm"},
        "kernels": {k: {"launches": v[0], "ms": round(v[1], 3), "GBps": round(v[2] / (v[1] * 1e-3) / 1e9, 1)}
                    for k, v in prof.items() if v[0] > 0},
    }
    ok = all(runs[k]["err"] < 1e-9 for k in ("warm", "plain", "warm_ref", "plain_ref"))
    if not ok:
        print("convergence check failed: |dF/F| " + ", ".join(f"{k} {runs[k]['err']:.2e}" for k in
                                                            ("warm", "plain", "warm_ref", "plain_ref")), file=sys.stderr)
    print(json.dumps(line), flush=True)
    if not ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
