"""The HIP path's pointwise White-Bear, Mark II (both forms) and tensor White-Bear code (energy density and the five
partials, deft_amd/csrc/wb_point.hpp) built for the host and checked against the
oracle on both sides of the series cut-offs (|n3| = 0.05), with no GPU."""

import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import fmt_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("point") / "point_harness")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950", "--cuda-host-only",
                    os.path.join(HERE, "host", "point_harness.cpp"), "-o", exe], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return exe


def _inputs():
    rng = np.random.default_rng(9)
    n3 = np.concatenate([rng.uniform(-0.049, 0.049, 40), rng.uniform(0.051, 0.6, 40),
                         [0.0, 1e-9, -1e-9, 0.0499999, 0.0500001, -0.0499999, -0.0500001]])
    m = len(n3)
    n2 = rng.uniform(0.0, 3.0, m)
    v = rng.uniform(-0.7, 0.7, (m, 3))
    return n3, n2, v


@pytest.mark.parametrize("functional,code", [("wb", 0), ("wb2", 1), ("wb2slow", 2)])
def test_point_code_matches_oracle(harness, functional, code):
    n3, n2, v = _inputs()
    text = "".join("%d %.17g %.17g %.17g %.17g %.17g\n" % (code, a, b, *c) for a, b, c in zip(n3, n2, v))
    out = subprocess.run([harness], input=text.encode(), stdout=subprocess.PIPE, check=True).stdout.decode()
    got = np.array([[float(x) for x in line.split()] for line in out.splitlines()])
    p1, p2, p3, P = O.whitebear_phi(n3, n2, v[:, 0], v[:, 1], v[:, 2], 1.0, functional=functional)
    _check(got, np.stack([p1, p2, p3, *P], axis=1))


def _check(got, ref):
    scale = np.maximum(np.abs(ref), 1e-3)
    nref = ref.shape[1]
    assert np.max(np.abs(got[:, :nref] - ref) / scale) < 1e-13
    # the energy-only function (what every kernel sums) agrees as well
    assert np.max(np.abs(got[:, nref:nref + 3] - ref[:, :3]) / scale[:, :3]) < 1e-13


def test_tensor_point_code_matches_oracle(harness):
    n3, n2, v = _inputs()
    t = np.random.default_rng(10).uniform(-0.4, 0.4, (len(n3), 6))
    text = "".join("3 " + " ".join("%.17g" % q for q in (a, b, *c, *d)) + "\n" for a, b, c, d in zip(n3, n2, v, t))
    out = subprocess.run([harness], input=text.encode(), stdout=subprocess.PIPE, check=True).stdout.decode()
    got = np.array([[float(x) for x in line.split()] for line in out.splitlines()])
    p1, p2, p3, P = O.whitebear_tensor_phi(n3, n2, v[:, 0], v[:, 1], v[:, 2], *t.T, 1.0)
    _check(got, np.stack([p1, p2, p3, *P], axis=1))


def test_log1m_matches_log1p(harness):
    """log1m(x) = log(1 - x), the kernels' replacement for log1p(-x): within 2 ulp of
    numpy's log1p over small, moderate and near-1 arguments, both sides of the k = 0
    region's edge |x| = 1 - sqrt(1/2), negative arguments, and the special values."""
    rng = np.random.default_rng(11)
    edge = 1 - np.sqrt(0.5)
    x = np.concatenate([
        rng.uniform(-0.05, 0.05, 400), rng.uniform(0.05, 0.999, 400), rng.uniform(-3.0, 0.0, 200),
        10.0 ** rng.uniform(-300, -1, 300), -(10.0 ** rng.uniform(-300, -1, 200)),
        1 - 10.0 ** rng.uniform(-16, -1, 200), -(10.0 ** rng.uniform(0, 300, 100)),
        np.nextafter(edge, 0) * np.array([1, -1]), np.nextafter(edge, 1) * np.array([1, -1]),
        [0.0, -0.0, 5e-324, 0.5, 0.75, 1 - 2.0 ** -53, np.nextafter(1.0, 0.0)]])
    text = "".join("-1 %.17g\n" % v for v in x)
    out = subprocess.run([harness], input=text.encode(), stdout=subprocess.PIPE, check=True).stdout.decode()
    got = np.array([float(v) for v in out.split()])
    want = np.log1p(-x)
    ulp = np.spacing(np.abs(want))
    assert np.max(np.abs(got - want) / np.maximum(ulp, 5e-324)) <= 2.0
    sp = "-1 1\n-1 1.5\n-1 nan\n-1 2\n"
    out = subprocess.run([harness], input=sp.encode(), stdout=subprocess.PIPE, check=True).stdout.decode().split()
    assert float(out[0]) == -np.inf and all(np.isnan(float(v)) for v in out[1:])
