"""Trains the warm-start network on converged densities from the library's own
minimizer, then measures what the warm start saves on held-out geometries.

    python -m deft_amd.train_warmstart --samples 800 --steps 3000 --out deft_amd/models/unet_warmstart.npz

Training data: random hard spheres, hard slabs and soft Gaussian potentials on a
periodic 64^3 grid of edge 6.4 (dx = 0.1, R = 1, kT = 1), bulk packing
fraction uniform in [0.1, 0.45], each converged to relative precision 1e-10.
The network is fully convolutional, so weights trained at 64^3 apply to any
cubic grid with the same spacing.
"""

import argparse
import json
import math
import sys
import time

import numpy as np

from . import WhiteBearFluidVeff
from . import nn as NN
from . import workloads as W


def log(msg):
    print(msg, flush=True)


def compare_starts(net, M, L, cases, relative_precision=1e-11):
    """Plain (bulk) start vs warm start on each (vext, eta): CG iterations, seconds, energies."""
    f = WhiteBearFluidVeff((M, M, M), np.diag([L] * 3), R=1.0, kT=1.0, mu=0.0)
    out = []
    for V, eta in cases:
        NN.solve(f, V, eta, relative_precision=relative_precision)  # first calls allocate
        NN.solve(f, V, eta, relative_precision=relative_precision, warm=net)
        t0 = time.perf_counter()
        _, sp = NN.solve(f, V, eta, relative_precision=relative_precision)
        tp = time.perf_counter() - t0
        t0 = time.perf_counter()
        _, sw = NN.solve(f, V, eta, relative_precision=relative_precision, warm=net)
        tw = time.perf_counter() - t0
        out.append({"eta": eta, "plain_iter": sp["iter"], "warm_iter": sw["iter"], "plain_s": tp, "warm_s": tw,
                    "plain_phases": [sp["fp32_iterations"], sp["mixed_iterations"]],
                    "warm_phases": [sw["fp32_iterations"], sw["mixed_iterations"]],
                    "dF_over_F": abs(sw["energy"] - sp["energy"]) / abs(sp["energy"]),
                    "plain_stop": sp["stop_reason"], "warm_stop": sw["stop_reason"]})
        log(f"  eta {eta:.3f}: plain {sp['iter']} it (fp32 {sp['fp32_iterations']}, mixed {sp['mixed_iterations']}) "
            f"{tp * 1e3:.1f} ms, warm {sw['iter']} it (fp32 {sw['fp32_iterations']}, mixed {sw['mixed_iterations']}) "
            f"{tw * 1e3:.1f} ms, |dF/F| {out[-1]['dF_over_F']:.2e}")
    f.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=800)
    ap.add_argument("--holdout", type=int, default=16)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--M", type=int, default=64)
    ap.add_argument("--L", type=float, default=6.4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=NN.WEIGHTS_PATH)
    ap.add_argument("--init", default=None, help="start from these weights instead of a fresh network")
    ap.add_argument("--report", default=None, help="write the evaluation as JSON here")
    args = ap.parse_args(argv)

    t0 = time.perf_counter()
    data = NN.make_dataset(args.samples, M=args.M, L=args.L, seed=args.seed, log=log)
    held = NN.make_dataset(args.holdout, M=args.M, L=args.L, seed=args.seed + 1000)
    log(f"datasets ready in {time.perf_counter() - t0:.1f} s")
    net = NN.WarmStartNet.load(args.init) if args.init else NN.WarmStartNet(seed=args.seed)
    rms0 = NN.evaluate(net, held)
    log(f"held-out RMS error of u before training: {np.mean(rms0):.4f}")
    hist = NN.train(net, data, args.steps, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
    rms = NN.evaluate(net, held)
    log(f"held-out RMS error of u after training: {np.mean(rms):.4f} (per sample {np.round(rms, 4).tolist()})")
    net.save(args.out, samples=args.samples, steps=args.steps, grid=args.M, cell=args.L,
             holdout_rms=float(np.mean(rms)), final_loss=float(np.mean(hist[-100:])))
    log(f"saved {args.out}")

    rng = np.random.default_rng(args.seed + 2000)
    log("held-out geometries (64^3):")
    small = compare_starts(net, args.M, args.L, [(NN.random_vext(args.M, args.L, rng), float(rng.uniform(0.15, 0.45)))
                                                 for _ in range(6)])
    log("radius-3 hard solute, eta 0.40, 128^3, 12.8^3 cell:")
    big = compare_starts(net, 128, 12.8, [(W.c2_solute_vext(128, 12.8, 3.0), 0.40)])
    report = {"holdout_rms_u": float(np.mean(rms)), "final_loss": float(np.mean(hist[-100:])),
              "seconds": time.perf_counter() - t0, "heldout_64": small, "solute_128": big}
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(report, fh, indent=1)
    return report


if __name__ == "__main__":
    main(sys.argv[1:])
