"""Extracts the reference's converged solute profiles into tests/golden/contact_inner_sphere.json.

    python tests/golden/make_contact_fixtures.py

Source: papers/contact/figs/inner-sphere{WB,WBm2}-04.0-{eta}.dat of the reference (and
their -mean.dat companions where published), the
output of papers/contact/figs/inner-sphere.cpp:177-249 (radial_plot, :143-172): a hard
solute of excluded-centre radius 2 (diameter argument 4, notinwall at :30-38) in a
periodic 20^3 cell (padding 16), grid 201^3 (GridDescription(lat, 0.1),
src/GridDescription.cpp:28), R = 1, kT = 1, mu from the bulk, Veff started at
-ln(n_b c + 1e-4 n_b), old-API preconditioned CG to |E_bulk V| 1e-10 for at most 200
iterations.  Column 0 is y of grid point (0, y, 0), column 1 the density there,
columns 2-7 the contact-density estimators written beside it (:220-233): n0,
gSigmaS, YuWu_S, nA, gSigmaA, Gross.
"WB" is HardSpheresNoTensor2 (White Bear), "WBT" TensorWhiteBear (files at eta 0.30
and 0.40 only), "WBm2" WhiteBearMarkII.
The columns are kept as printed (%g).
"""

import json
import os

REF = "/root/reference/papers/contact/figs"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    out = {"diameter": 4.0, "cell": 20.0, "n": 201, "R": 1.0, "kT": 1.0, "profiles": {}}
    for name in ("WB", "WBT", "WBm2"):
        for eta in ("0.10", "0.20", "0.30", "0.40", "0.50"):
            path = os.path.join(REF, "inner-sphere%s-04.0-%s.dat" % (name, eta))
            if not os.path.exists(path):
                continue
            rows = []
            with open(path) as fh:
                for line in fh:
                    cols = line.split()
                    if len(cols) < 8:
                        continue
                    rows.append([float(c) for c in cols[:8]])
            prof = {"eta": float(eta), "y": [r[0] for r in rows], "density": [r[1] for r in rows]}
            for c, key in enumerate(("n0", "gSigmaS", "YuWu_S", "nA", "gSigmaA", "Gross")):
                prof[key] = [r[2 + c] for r in rows]
            out["profiles"]["%s-%s" % (name, eta)] = prof
            # radial_plot2's companion (:114-147, written at :250-252): the same seven
            # fields averaged over twenty directions at radii 0, 0.1, ... (first column r)
            mean = os.path.join(REF, "inner-sphere%s-04.0-%s-mean.dat" % (name, eta))
            if os.path.exists(mean):
                rows = []
                with open(mean) as fh:
                    for line in fh:
                        cols = line.split()
                        if len(cols) >= 8:
                            rows.append([float(c) for c in cols[:8]])
                m = {"r": [r[0] for r in rows], "density": [r[1] for r in rows]}
                for c, key in enumerate(("n0", "gSigmaS", "YuWu_S", "nA", "gSigmaA", "Gross")):
                    m[key] = [r[2 + c] for r in rows]
                out.setdefault("means", {})["%s-%s" % (name, eta)] = m
    with open(os.path.join(HERE, "contact_inner_sphere.json"), "w") as fh:
        json.dump(out, fh, indent=0)


if __name__ == "__main__":
    main()
