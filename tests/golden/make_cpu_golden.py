"""Runs the oracle's C restatement (oracle/build/fmt_cpu, built by `make -C oracle`) on
the 128^3 solute workload to convergence under the GPU bench's rule and writes
tests/golden/cpu_c2_rule.json: the CPU path's own iteration and evaluation counts,
which bench.py's cpu_baseline multiplies by per-evaluation times sampled on the host.

    python tests/golden/make_cpu_golden.py

Workload: R = kT = 1, eta = 0.40, hard solute of excluded-centre radius 3 at the origin
of a periodic 12.8^3 cell on 128^3 points, mu from the bulk, Veff = -ln n_b, the new
API's preconditioned Minimize (FP64 throughout) to absolute precision |p V| 1e-10
(papers/contact/figs/inner-sphere.cpp:201-203).
"""

import json
import math
import os
import platform
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import fmt_oracle as O  # noqa: E402


def main():
    nb = O.bulk_density(0.40, 1.0)
    precision = abs(O.pressure(nb, 1.0) * 12.8 ** 3) * 1e-10
    exe = os.path.join(ROOT, "oracle", "build", "fmt_cpu")
    out = subprocess.run([exe, "solute", "128", "12.8", "0.40", "3.0", repr(precision), "0", "0"],
                         capture_output=True, text=True, check=True).stdout
    r = json.loads(out)
    r["precision"] = precision
    r["host"] = platform.processor() or platform.machine()
    gold = json.load(open(os.path.join(HERE, "oracle_c2.json")))
    r["dF_over_F_vs_oracle_c2"] = abs(r["energy"] - gold["energy"]) / abs(gold["energy"])
    with open(os.path.join(HERE, "cpu_c2_rule.json"), "w") as fh:
        json.dump(r, fh, indent=1)
    print(r)


if __name__ == "__main__":
    main()
