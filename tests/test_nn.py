"""Warm-start network: the float64 oracle pinned by finite differences and exact
bfloat16 rounding cases (CPU), then the device forward, backward, training and
warm start against it (GPU)."""

import math
import os

import numpy as np
import pytest

import deft_amd
import deft_amd.nn
from oracle import unet_oracle as U

SMALL = (16, 16, 16, 16)


def random_params(channels, seed):
    convs, hw, hb, n = U.layout(channels)
    rng = np.random.default_rng(seed)
    p = np.zeros(n)
    for c in convs:
        W = rng.uniform(-1, 1, (c["cout"], c["kp"])) * math.sqrt(6.0 / (27 * c["cin"]))
        W[:, 27 * c["cin"]:] = 0
        if c["woff"] == 0:  # input channels 2..7 carry zeros
            k = np.arange(c["kp"])
            W[:, (k < 27 * 8) & (k % 8 >= 2)] = 0
        p[c["woff"]:c["woff"] + W.size] = W.ravel()
        p[c["boff"]:c["boff"] + c["cout"]] = rng.uniform(0.0, 0.1, c["cout"])
    p[hw:hb] = rng.uniform(-0.5, 0.5, channels[0])
    p[hb] = 0.1
    return p.astype(np.float32)


def random_inputs(B, M, seed):
    rng = np.random.default_rng(seed)
    boltz = np.where(rng.random((B, M, M, M)) > 0.2, rng.random((B, M, M, M)), 0.0).astype(np.float32)
    eta = rng.uniform(0.1, 0.45, B).astype(np.float32)
    target = rng.normal(size=(B, M, M, M)).astype(np.float32)
    mask = (boltz > 0).astype(np.float32)
    return boltz, eta, target, mask


# ----------------------------------------------------------------- CPU: oracle

def test_bf16_rounding_known_values():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -2.5, 1 + 2 ** -8 + 2 ** -20, 3.0e38],
                 dtype=np.float32)
    # ties go to the even mantissa; anything past the tie rounds up
    want = [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 1 + 2 ** -7, 2.9979e38]
    got = U.bf16(x)
    assert np.allclose(got[:6], want[:6], rtol=0, atol=0)
    assert abs(got[6] / 3.0e38 - 1) < 2 ** -8


def test_layout_counts():
    convs, hw, hb, n = U.layout((16, 32, 48, 64))
    assert [c["kp"] for c in convs[:4]] == [224, 448, 448, 864]
    assert n == sum(c["cout"] * c["kp"] + c["cout"] for c in convs) + 16 + 1
    assert convs[1]["woff"] == 16 * 224 + 16


def test_pool_up_are_adjoint():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 8, 8, 8, 3))
    y = rng.normal(size=(2, 4, 4, 4, 3))
    assert abs((U.pool(x) * y).sum() - (x * U.pool_backward(y)).sum()) < 1e-12
    assert abs((U.up(y) * x).sum() - (y * U.up_backward(x)).sum()) < 1e-12
    assert np.allclose(U.pool(U.up(y)), y)


def test_conv_is_a_periodic_correlation():
    """A single tap of the convolution shifts the input by that tap's offset."""
    c = dict(cin=8, cout=16, kp=224, woff=0, boff=16 * 224)
    p = np.zeros(16 * 224 + 16)
    t = 9 * 2 + 3 * 0 + 1  # offset (+1, -1, 0)
    p[3 * 224 + t * 8 + 0] = 1.0  # out channel 3 reads input channel 0 at that tap
    x = np.zeros((1, 8, 8, 8, 8))
    x[0, 5, 0, 4, 0] = 2.0
    out = U.conv(x, p, c, False)
    assert out[0, 4, 1, 4, 3] == 2.0  # v + (1, -1, 0) = (5, 0, 4)
    assert out.sum() == 2.0


def test_oracle_gradient_matches_finite_differences():
    ch = SMALL
    p = random_params(ch, 1).astype(np.float64)
    boltz, eta, target, mask = random_inputs(2, 16, 2)
    L0, g = U.loss_and_grad(p, boltz, eta, target, mask, ch)
    rng = np.random.default_rng(3)
    convs, hw, hb, n = U.layout(ch)
    blocks = [(c["woff"], c["woff"] + c["cout"] * c["kp"]) for c in convs]
    blocks += [(c["boff"], c["boff"] + c["cout"]) for c in convs] + [(hw, hb + 1)]
    # central differences along a random direction inside each parameter block; the
    # step is small enough that no ReLU input changes sign
    e = 1e-7
    for a, b in blocks:
        d = np.zeros(n)
        d[a:b] = rng.normal(size=b - a)
        for c in convs:  # keep the zero padding zero
            d[c["woff"]:c["woff"] + c["cout"] * c["kp"]].reshape(c["cout"], c["kp"])[:, 27 * c["cin"]:] = 0
        fd = (U.loss_and_grad(p + e * d, boltz, eta, target, mask, ch)[0]
              - U.loss_and_grad(p - e * d, boltz, eta, target, mask, ch)[0]) / (2 * e)
        assert abs(fd - g @ d) < 1e-6 * abs(g @ d), (a, b, fd, g @ d)


def test_random_vext_shapes():
    rng = np.random.default_rng(0)
    seen_inf = False
    for _ in range(12):
        V = deft_amd.nn.random_vext(32, 3.2, rng)
        assert V.shape == (32, 32, 32)
        assert np.isfinite(V).mean() > 0.3
        assert not np.any(np.isnan(V))
        seen_inf |= bool(np.isinf(V).any())
    assert seen_inf


def oracle_envelope(p, boltz, eta, channels, want, seed=0):
    """RMS change of the BF16-emulated oracle output under a 1e-6 relative perturbation
    of the weights: how far FP32 accumulation-order differences alone can move u for
    this network (BF16 rounding flips cascade through the layers of some networks)."""
    rng = np.random.default_rng(seed)
    q = np.asarray(p, dtype=np.float64) * (1 + 1e-6 * rng.standard_normal(len(p)))
    pert, _ = U.forward(q, boltz, eta, channels, emulate_bf16=True)
    return np.sqrt(np.mean((pert - want) ** 2))


def conv_case(rng, B, M, cin, cout):
    x = np.maximum(rng.normal(size=(B, M, M, M, cin)), 0).astype(np.float32)
    kp = (27 * cin + 31) // 32 * 32
    w = np.zeros((cout, kp), dtype=np.float32)
    w[:, :27 * cin] = rng.normal(size=(cout, 27 * cin)) / math.sqrt(27 * cin)
    bias = rng.normal(size=cout).astype(np.float32) * 0.1
    c = dict(cin=cin, cout=cout, kp=kp, woff=0, boff=cout * kp)
    params = np.concatenate([w.ravel(), bias])
    return x, w, bias, c, params


# ----------------------------------------------------------------- GPU

def _dev(a):
    return deft_amd.nn.DeviceArray.from_numpy(a)


def transform_numpy(x, t):
    """deft_net_make_batch's mapping restated: output axis a reads input axis t[a] at
    flip ? M-1-((o-s) mod M) : (o-s) mod M."""
    M = x.shape[0]
    o = np.indices((M, M, M))
    src = [None] * 3
    for a in range(3):
        c = (o[a] - t[6 + a]) % M
        src[t[a]] = M - 1 - c if t[3 + a] else c
    return x[src[0], src[1], src[2]]


@pytest.mark.gpu
def test_mfma_lane_layout(gpu):
    rng = np.random.default_rng(5)
    a = rng.normal(size=(16, 32)).astype(np.float32)
    b = rng.normal(size=(16, 32)).astype(np.float32)
    c = deft_amd.nn.mfma_probe(a, b)
    want = U.bf16(a) @ U.bf16(b).T
    assert np.max(np.abs(c - want)) < 1e-5 * np.max(np.abs(want))


CONV_CASES = [  # (B, M, cin, cout, tiled): tiled applies where M % 16 == 0
    (2, 16, 8, 16, True), (1, 16, 16, 32, True), (1, 16, 16, 32, False), (1, 32, 48, 48, True),
    (1, 16, 64, 48, True), (2, 8, 64, 64, False), (1, 16, 32, 64, True), (2, 4, 32, 16, False),
    (2, 2, 16, 16, False)]


@pytest.mark.gpu
@pytest.mark.parametrize("B,M,cin,cout,tiled", CONV_CASES)
def test_conv_layer_matches_oracle(gpu, B, M, cin, cout, tiled):
    """One convolution (forward, backward-data, weight gradient) against the oracle's with
    the same BF16 operand rounding: only FP32 accumulation order differs."""
    rng = np.random.default_rng(B * 1000 + M * 10 + cin + cout)
    x, w, bias, c, params = conv_case(rng, B, M, cin, cout)
    for relu in (False, True):
        got = deft_amd.nn.layer_probe("conv", x, cout, w, bias, relu=relu, tiled=tiled)
        W = U._weights(params, c, True)
        pre = np.zeros(x.shape[:4] + (cout,))
        mag = np.zeros_like(pre)
        xb = U.bf16(x)
        for t in range(27):
            dx, dy, dz = U._tap(t)
            xs = np.roll(xb, (-dx, -dy, -dz), axis=(1, 2, 3))
            pre += xs @ W[:, t, :].T
            mag += np.abs(xs) @ np.abs(W[:, t, :]).T
        want = pre + bias if not relu else np.maximum(pre + bias, 0)
        assert np.all(np.abs(got - want) <= 1e-6 * (mag + 1)), (relu, np.max(np.abs(got - want)))
    d = rng.normal(size=(B, M, M, M, cout)).astype(np.float32)
    gW, _, din = U.conv_backward(d, np.ones_like(d), x, params, c, True)
    got_din = deft_amd.nn.layer_probe("dgrad", d, cin, w, tiled=tiled) if cin % 16 == 0 else None
    if got_din is not None:
        assert np.max(np.abs(got_din - din)) <= 1e-5 * np.max(np.abs(din))
    got_gw = deft_amd.nn.layer_probe("wgrad", x, x2=d)
    assert np.max(np.abs(got_gw - gW)) <= 1e-5 * np.max(np.abs(gW))
    assert not np.any(got_gw[:, 27 * cin:])


@pytest.mark.gpu
@pytest.mark.parametrize("M", [4, 16])
def test_pool_and_upsample_layers(gpu, M):
    rng = np.random.default_rng(M)
    x = rng.normal(size=(2, M, M, M, 16)).astype(np.float32)
    got = deft_amd.nn.layer_probe("pool", x)
    assert np.allclose(got, U.pool(x.astype(np.float64)), rtol=0, atol=1e-6)
    y = rng.normal(size=(2, M // 2, M // 2, M // 2, 32)).astype(np.float32)
    assert np.array_equal(deft_amd.nn.layer_probe("up", y), U.up(y).astype(np.float32))


@pytest.mark.gpu
def test_layout_matches_oracle(gpu):
    net = deft_amd.nn.WarmStartNet()
    convs, hw, hb, n = U.layout(net.channels)
    assert net.num_params == n
    for i, c in enumerate(convs):
        assert net.layout(i) == (c["cin"], c["cout"], c["kp"], c["woff"], c["boff"])
    assert net.layout(-1)[3:] == (hw, hb)
    p = net.params
    for c in convs:  # padding of the initial weights is zero
        W = p[c["woff"]:c["woff"] + c["cout"] * c["kp"]].reshape(c["cout"], c["kp"])
        assert not np.any(W[:, 27 * c["cin"]:])
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("channels", [SMALL, (16, 32, 48, 64)])
def test_forward_matches_oracle(gpu, channels):
    net = deft_amd.nn.WarmStartNet(channels)
    p = random_params(channels, 7)
    net.params = p
    boltz, eta, _, _ = random_inputs(2, 16, 8)
    u = net.forward(_dev(boltz), eta).numpy()
    want, _ = U.forward(p, boltz, eta, channels, emulate_bf16=True)
    # FP32 accumulation order now and then moves an activation across a BF16 rounding
    # boundary (2^-8 relative), and such flips propagate: the device must stay within
    # the oracle's own envelope for that effect (each layer is checked exactly above).
    env = oracle_envelope(p, boltz, eta, channels, want)
    rms = np.sqrt(np.mean((u - want) ** 2))
    assert rms < 3 * env + 1e-4 * np.sqrt(np.mean(want ** 2)), (rms, env)
    net.close()


@pytest.mark.gpu
def test_loss_grad_matches_oracle(gpu):
    ch = (16, 32, 48, 64)
    net = deft_amd.nn.WarmStartNet(ch)
    p = random_params(ch, 11)
    net.params = p
    boltz, eta, target, mask = random_inputs(2, 16, 12)
    loss, g = net.loss_grad(_dev(boltz), eta, _dev(target), _dev(mask))
    lo, go = U.loss_and_grad(p, boltz, eta, target, mask, ch, emulate_bf16=True)
    assert abs(loss - lo) < 5e-3 * lo
    convs, hw, hb, n = U.layout(ch)
    blocks = [(c["woff"], c["woff"] + c["cout"] * c["kp"]) for c in convs]
    blocks += [(c["boff"], c["boff"] + c["cout"]) for c in convs] + [(hw, hb), (hb, hb + 1)]
    for a, b in blocks:
        ref = go[a:b]
        assert np.linalg.norm(g[a:b] - ref) <= 2e-2 * np.linalg.norm(ref) + 1e-7, (a, b)
    net.close()


@pytest.mark.gpu
def test_tiled_and_untiled_convolutions_agree(gpu):
    """M = 32: levels 0 and 1 take the LDS-tiled kernel by default; both kernels form
    the same BF16 products and accumulate them in the same order."""
    ch = (16, 32, 48, 64)
    net = deft_amd.nn.WarmStartNet(ch)
    net.params = random_params(ch, 21)
    boltz, eta, target, _ = random_inputs(2, 32, 22)
    b, t = _dev(boltz), _dev(target)
    u_tiled = net.forward(b, eta).numpy()
    l_tiled, g_tiled = net.loss_grad(b, eta, t)
    net.set_conv_kernel(False)
    u_plain = net.forward(b, eta).numpy()
    l_plain, g_plain = net.loss_grad(b, eta, t)
    assert np.array_equal(u_tiled, u_plain)
    assert abs(l_tiled - l_plain) < 1e-6 * l_plain
    assert np.linalg.norm(g_tiled - g_plain) < 1e-5 * np.linalg.norm(g_plain)
    net.close()


@pytest.mark.gpu
def test_fused_inference_matches_layer_by_layer_forward(gpu):
    """At 128^3 (every level a multiple of 16) inference folds the upsampling, the skip
    sums and the head into the convolutions; the untiled path runs layer by layer.
    Only the head's summation order differs."""
    net = deft_amd.nn.WarmStartNet.load() if os.path.exists(deft_amd.nn.WEIGHTS_PATH) else \
        deft_amd.nn.WarmStartNet()
    rng = np.random.default_rng(31)
    M = 128
    V = deft_amd.nn.random_vext(M, 12.8, rng)
    ok = np.isfinite(V)
    b = _dev(np.where(ok, np.exp(-np.where(ok, V, 0)), 0)[None])
    u_fused = net.forward(b, [0.35]).numpy()
    net.set_conv_kernel(False)
    u_plain = net.forward(b, [0.35]).numpy()
    assert np.max(np.abs(u_fused - u_plain)) <= 1e-5 * max(1.0, np.max(np.abs(u_plain)))
    net.close()


@pytest.mark.gpu
def test_loss_without_mask_uses_boltz(gpu):
    net = deft_amd.nn.WarmStartNet(SMALL, seed=2)
    boltz, eta, target, mask = random_inputs(2, 16, 13)
    b, t = _dev(boltz), _dev(target)
    l1, g1 = net.loss_grad(b, eta, t, _dev(mask))
    l2, g2 = net.loss_grad(b, eta, t)
    # the weight gradients are summed with atomics, so agreement is to FP32 rounding
    assert abs(l1 - l2) < 1e-6 * l1
    assert np.linalg.norm(g1 - g2) < 1e-5 * np.linalg.norm(g1)
    net.close()


@pytest.mark.gpu
def test_make_batch_matches_numpy(gpu):
    rng = np.random.default_rng(6)
    M, n = 16, 3
    data = rng.normal(size=(n, M, M, M)).astype(np.float32)
    d = _dev(data)
    samples = np.array([2, 0, 2, 1], dtype=np.int32)
    xf = deft_amd.nn.random_transforms(rng, 4, M)
    xf[0] = [0, 1, 2, 0, 0, 0, 0, 0, 0]  # identity
    out = deft_amd.nn.DeviceArray((4, M, M, M))
    deft_amd.nn.make_batch(d, n, M, samples, xf, out)
    got = out.numpy()
    assert np.array_equal(got[0], data[2])
    for b in range(4):
        assert np.array_equal(got[b], transform_numpy(data[samples[b]], xf[b])), b
    with pytest.raises(deft_amd.DeftError):
        deft_amd.nn.make_batch(d, n, M, [3], xf[:1], out)
    bad = xf[:1].copy()
    bad[0, :3] = [0, 0, 2]
    with pytest.raises(deft_amd.DeftError):
        deft_amd.nn.make_batch(d, n, M, [0], bad, out)


@pytest.mark.gpu
def test_training_steps_reduce_loss(gpu):
    net = deft_amd.nn.WarmStartNet(SMALL, seed=3)
    boltz, eta, _, mask = random_inputs(2, 16, 4)
    # a smooth target the network can fit: a function of the local Boltzmann factor
    target = np.where(mask > 0, 0.5 * boltz - 0.2, 0).astype(np.float32)
    b, t = _dev(boltz), _dev(target)
    first = net.train_step(b, eta, t, 3e-3)
    for _ in range(100):
        last = net.train_step(b, eta, t, 3e-3)
    assert last < 0.5 * first, (first, last)
    assert np.all(np.isfinite(net.forward(b, eta).numpy()))
    net.close()


@pytest.mark.gpu
def test_warm_start_rejects_bad_plans(gpu):
    net = deft_amd.nn.WarmStartNet(SMALL)
    f = deft_amd.WhiteBearFluidVeff((32, 32, 20), np.diag([3.2, 3.2, 2.0]), R=1, kT=1, mu=0.0)
    with pytest.raises(deft_amd.DeftError):
        net.warm_start(f, 0.3)
    g = deft_amd.WhiteBearFluidVeff((32, 32, 32), np.diag([3.2] * 3), R=1, kT=1, mu=0.0)
    with pytest.raises(deft_amd.DeftError):
        net.warm_start(g, 0.9)
    f.close()
    g.close()
    net.close()


@pytest.mark.gpu
def test_warm_start_sets_veff_from_prediction(gpu):
    net = deft_amd.nn.WarmStartNet(SMALL, seed=9)
    M, L, eta = 32, 3.2, 0.3
    V = deft_amd.nn.random_vext(M, L, np.random.default_rng(1))
    f = deft_amd.WhiteBearFluidVeff((M, M, M), np.diag([L] * 3), R=1, kT=1.0, mu=0.0)
    f.Vext = V
    lam = net.warm_start(f, eta)
    ok = np.isfinite(V)
    boltz = np.where(ok, np.exp(-np.where(ok, V, 0)), 0).astype(np.float32)
    u = net.forward(_dev(boltz[None]), [eta]).numpy()[0]
    nb = eta / (4 * math.pi / 3)
    want = -(math.log(nb) + lam * u.astype(np.float64))
    assert np.max(np.abs(f.Veff - want)) < 1e-4 * np.max(np.abs(want))
    f.close()
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("factor", [1, 2])
def test_batched_warm_start_equals_one_by_one(gpu, factor):
    """deft_plans_warm_start: one inference over several plans gives each plan the Veff
    (and blend factor) its own warm start gives, including a plan whose prediction has to
    be pulled toward the bulk start; mixed etas, both grid factors."""
    net = deft_amd.nn.WarmStartNet(SMALL, seed=5)
    M, L = 32 * factor, 3.2
    rng = np.random.default_rng(21)
    cases = [(deft_amd.nn.random_vext(M, L, rng), eta) for eta in (0.15, 0.3, 0.42)]
    fl = []
    for V, eta in cases:
        f = deft_amd.WhiteBearFluidVeff((M, M, M), np.diag([L] * 3), R=1, kT=1.0,
                                        mu=deft_amd.find_chemical_potential(eta / (4 * math.pi / 3)))
        f.Vext = V
        fl.append(f)
    alone = []
    for f, (V, eta) in zip(fl, cases):
        lam = net.warm_start(f, eta, factor=factor)
        alone.append((lam, f.Veff.copy()))
    lams = deft_amd.nn.warm_start_batch(net, fl, [eta for _, eta in cases], factor=factor)
    for f, lam, (lam1, v1) in zip(fl, lams, alone):
        assert lam == lam1
        v = f.Veff
        fin = np.isfinite(v1)
        assert np.array_equal(np.isfinite(v), fin)
        assert np.max(np.abs(v[fin] - v1[fin])) <= 1e-6 * np.max(np.abs(v1[fin]))
        f.close()
    net.close()


@pytest.mark.gpu
def test_scaled_warm_start_matches_numpy(gpu):
    """factor 2: the network runs on the 2x coarser grid (tent-smoothed Boltzmann factor)
    and its u is interpolated back trilinearly."""
    net = deft_amd.nn.WarmStartNet(SMALL, seed=4)
    M, L, eta = 64, 3.2, 0.3
    rng = np.random.default_rng(11)
    V = deft_amd.nn.random_vext(M, L, rng)
    f = deft_amd.WhiteBearFluidVeff((M, M, M), np.diag([L] * 3), R=1, kT=1.0,
                                    mu=deft_amd.find_chemical_potential(eta / (4 * math.pi / 3)))
    f.Vext = V
    lam = net.warm_start(f, eta, factor=2)
    ok = np.isfinite(V)
    b = np.where(ok, np.exp(-np.where(ok, V, 0)), 0).astype(np.float64)
    for ax in range(3):
        b = 0.25 * np.roll(b, 1, axis=ax) + 0.5 * b + 0.25 * np.roll(b, -1, axis=ax)
    bc = b[::2, ::2, ::2].astype(np.float32)
    u = net.forward(_dev(bc[None]), [eta]).numpy()[0].astype(np.float64)
    for ax in range(3):
        t = (np.arange(M) % 2) / 2.0
        shape = [1, 1, 1]
        shape[ax] = M
        t = t.reshape(shape)
        lo = np.repeat(u, 2, axis=ax)
        hi = np.repeat(np.roll(u, -1, axis=ax), 2, axis=ax)
        u = (1 - t) * lo + t * hi
    want = -(math.log(eta / (4 * math.pi / 3)) + lam * u)
    got = f.Veff
    assert np.sqrt(np.mean((got - want) ** 2)) < 1e-3 * np.sqrt(np.mean(want ** 2))
    with pytest.raises(deft_amd.DeftError):
        net.warm_start(f, eta, factor=3)
    f.close()
    net.close()


@pytest.mark.gpu
def test_warm_start_keeps_the_start_energy_finite(gpu):
    """A prediction that packs the fluid past n3 = 1 (here u = 10 everywhere) is pulled
    toward the bulk start until the energy is finite."""
    net = deft_amd.nn.WarmStartNet(SMALL, seed=1)
    convs, hw, hb, n = U.layout(SMALL)
    p = net.params
    p[hw:hb] = 0
    p[hb] = 10.0
    net.params = p
    M, L, eta = 32, 3.2, 0.3
    f = deft_amd.WhiteBearFluidVeff((M, M, M), np.diag([L] * 3), R=1, kT=1.0,
                                    mu=deft_amd.find_chemical_potential(eta / (4 * math.pi / 3)))
    lam = net.warm_start(f, eta)
    assert 0 <= lam < 1
    assert np.isfinite(f.energy())
    # lambda u must keep n3 = 4 pi/3 n_b exp(lambda u) < 1
    assert eta * math.exp(10.0 * lam) < 1
    f.close()
    net.close()


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(deft_amd.nn.WEIGHTS_PATH), reason="no trained weights committed")
def test_trained_warm_start_saves_iterations(gpu):
    """Held-out geometries (a generator seed not used in training): the committed
    network's start converges to the same free energy as the bulk start, in fewer CG
    iterations overall and fewer FP32 iterations."""
    net = deft_amd.nn.WarmStartNet.load()
    M, L = 64, 6.4
    f = deft_amd.WhiteBearFluidVeff((M, M, M), np.diag([L] * 3), R=1, kT=1.0, mu=0.0)
    rng = np.random.default_rng(2024)
    plain_it = warm_it = plain_fp32 = warm_fp32 = 0
    for eta in (0.2, 0.3, 0.35, 0.42):
        V = deft_amd.nn.training_vext(M, L, rng)
        _, sp = deft_amd.nn.solve(f, V, eta, relative_precision=1e-11)
        _, sw = deft_amd.nn.solve(f, V, eta, relative_precision=1e-11, warm=net)
        assert sp["stop_reason"] == sw["stop_reason"] == "converged"
        assert abs(sw["energy"] - sp["energy"]) < 1e-9 * abs(sp["energy"]), eta
        plain_it += sp["iter"]
        warm_it += sw["iter"]
        plain_fp32 += sp["fp32_iterations"]
        warm_fp32 += sw["fp32_iterations"]
    assert warm_it < plain_it, (warm_it, plain_it)
    assert warm_fp32 < plain_fp32, (warm_fp32, plain_fp32)
    f.close()
    net.close()
