"""TEST INFRASTRUCTURE ONLY -- the CPU restatement of the warm-start network.

The product (deft_amd/csrc/nn.hip) must never import or call this module; tests
use it as the checker of the device forward and backward passes.

There is no counterpart of this network in deft (deft starts minimizations from
a bulk or hand-built Veff, e.g. papers/contact/figs/inner-sphere.cpp), so this
oracle restates the network definition itself:

  x0 = [exp(-Vext/kT), eta, 0 x 6]                   (8 input channels)
  encoder  level l: a = relu(conv(in)), s_l = relu(conv(a)), in_{l+1} = avgpool2(s_l)
  bottleneck:       b = relu(conv(relu(conv(in_3))))
  decoder  level l: h = relu(conv(up2(d_{l+1}))) + s_l, d_l = relu(conv(h))
  head:             u = bh + sum_c wh[c] d_0[c]

with periodic 3x3x3 convolutions
  out[v][co] = bias[co] + sum_{tap, ci} in[v + off(tap)][ci] W[co][tap * cin + ci],
  off(tap) = (tap // 9 - 1, tap // 3 % 3 - 1, tap % 3 - 1).

Parameter layout (one FP32 vector): for each of the 14 convolutions in order
W (cout x kp, kp = 27 cin rounded up to 32, zero padding) then its bias, then
the head weights (C0) and the head bias.

`emulate_bf16=True` rounds every convolution operand to bfloat16 (round to
nearest even), as the device's MFMA path does, so that device and oracle agree
to FP32 accumulation-order error.  The backward pass is derived by hand; its
float64 form (emulate_bf16=False) is pinned by central finite differences in
tests/test_nn.py.
"""

import numpy as np

DEFAULT_CHANNELS = (16, 32, 48, 64)


def conv_specs(channels=DEFAULT_CHANNELS):
    """(cin, cout, level) of the 14 convolutions, in parameter order."""
    C = channels
    return [(8, C[0], 0), (C[0], C[0], 0), (C[0], C[1], 1), (C[1], C[1], 1), (C[1], C[2], 2), (C[2], C[2], 2),
            (C[2], C[3], 3), (C[3], C[3], 3), (C[3], C[2], 2), (C[2], C[2], 2), (C[2], C[1], 1), (C[1], C[1], 1),
            (C[1], C[0], 0), (C[0], C[0], 0)]


def layout(channels=DEFAULT_CHANNELS):
    """List of dicts (cin, cout, kp, woff, boff) per convolution, head offsets, total size."""
    convs = []
    off = 0
    for cin, cout, level in conv_specs(channels):
        kp = (27 * cin + 31) // 32 * 32
        convs.append(dict(cin=cin, cout=cout, kp=kp, level=level, woff=off, boff=off + cout * kp))
        off += cout * kp + cout
    head_w = off
    head_b = off + channels[0]
    return convs, head_w, head_b, head_b + 1


def bf16(x):
    """Round to bfloat16 (nearest, ties to even) and return as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _tap(t):
    return t // 9 - 1, t // 3 % 3 - 1, t % 3 - 1


def _weights(params, c, emulate):
    W = np.asarray(params[c["woff"]:c["woff"] + c["cout"] * c["kp"]], dtype=np.float64)
    W = W.reshape(c["cout"], c["kp"])[:, :27 * c["cin"]].reshape(c["cout"], 27, c["cin"])
    return bf16(W) if emulate else W


def conv(x, params, c, emulate):
    """relu(conv(x) + bias); x has shape (B, M, M, M, cin)."""
    W = _weights(params, c, emulate)
    xb = bf16(x) if emulate else x
    out = np.zeros(x.shape[:4] + (c["cout"],))
    for t in range(27):
        dx, dy, dz = _tap(t)
        xs = np.roll(xb, (-dx, -dy, -dz), axis=(1, 2, 3))
        out += xs @ W[:, t, :].T
    out += np.asarray(params[c["boff"]:c["boff"] + c["cout"]], dtype=np.float64)
    return np.maximum(out, 0.0)


def conv_backward(d, act, x, params, c, emulate, need_input_grad=True):
    """Gradients of relu(conv(x)) given d = dL/d(output) and act = output."""
    W = _weights(params, c, emulate)
    dpre = np.where(act > 0, d, 0.0)
    gb = dpre.sum(axis=(0, 1, 2, 3))
    dp = bf16(dpre) if emulate else dpre
    xb = bf16(x) if emulate else x
    gW = np.zeros((c["cout"], c["kp"]))
    din = np.zeros_like(x) if need_input_grad else None
    d2 = dp.reshape(-1, c["cout"])
    for t in range(27):
        dx, dy, dz = _tap(t)
        xs = np.roll(xb, (-dx, -dy, -dz), axis=(1, 2, 3))
        gW[:, t * c["cin"]:(t + 1) * c["cin"]] = d2.T @ xs.reshape(-1, c["cin"])
        if need_input_grad:
            din += np.roll(dp @ W[:, t, :], (dx, dy, dz), axis=(1, 2, 3))
    return gW, gb, din


def pool(x):
    B, M = x.shape[0], x.shape[1]
    m = M // 2
    return x.reshape(B, m, 2, m, 2, m, 2, -1).mean(axis=(2, 4, 6))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, axis=1), 2, axis=2), 2, axis=3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, axis=1), 2, axis=2), 2, axis=3)


def up_backward(g):
    B, M = g.shape[0], g.shape[1]
    m = M // 2
    return g.reshape(B, m, 2, m, 2, m, 2, -1).sum(axis=(2, 4, 6))


def pack_input(boltz, eta):
    boltz = np.asarray(boltz, dtype=np.float64)
    x0 = np.zeros(boltz.shape + (8,))
    x0[..., 0] = boltz
    x0[..., 1] = np.asarray(eta, dtype=np.float64)[:, None, None, None]
    return x0


def forward(params, boltz, eta, channels=DEFAULT_CHANNELS, emulate_bf16=False):
    """u of shape (B, M, M, M) and the activations the backward pass needs."""
    convs, head_w, head_b, _ = layout(channels)
    p = np.asarray(params, dtype=np.float64)
    a = {"x0": pack_input(boltz, eta)}
    e = emulate_bf16
    a["a0"] = conv(a["x0"], p, convs[0], e)
    a["s0"] = conv(a["a0"], p, convs[1], e)
    a["p1"] = pool(a["s0"])
    a["a1"] = conv(a["p1"], p, convs[2], e)
    a["s1"] = conv(a["a1"], p, convs[3], e)
    a["p2"] = pool(a["s1"])
    a["a2"] = conv(a["p2"], p, convs[4], e)
    a["s2"] = conv(a["a2"], p, convs[5], e)
    a["p3"] = pool(a["s2"])
    a["a3"] = conv(a["p3"], p, convs[6], e)
    a["b3"] = conv(a["a3"], p, convs[7], e)
    a["u2"] = up(a["b3"])
    a["r2"] = conv(a["u2"], p, convs[8], e)
    a["h2"] = a["r2"] + a["s2"]
    a["d2"] = conv(a["h2"], p, convs[9], e)
    a["u1"] = up(a["d2"])
    a["r1"] = conv(a["u1"], p, convs[10], e)
    a["h1"] = a["r1"] + a["s1"]
    a["d1"] = conv(a["h1"], p, convs[11], e)
    a["u0"] = up(a["d1"])
    a["r0"] = conv(a["u0"], p, convs[12], e)
    a["h0"] = a["r0"] + a["s0"]
    a["d0"] = conv(a["h0"], p, convs[13], e)
    u = a["d0"] @ p[head_w:head_b] + p[head_b]
    return u, a


def loss_and_grad(params, boltz, eta, target, mask, channels=DEFAULT_CHANNELS, emulate_bf16=False):
    """Masked mean squared error of u against target and its gradient (parameter layout)."""
    convs, head_w, head_b, n = layout(channels)
    p = np.asarray(params, dtype=np.float64)
    u, a = forward(p, boltz, eta, channels, emulate_bf16)
    m = np.asarray(mask) > 0
    count = max(int(m.sum()), 1)
    err = np.where(m, u - np.asarray(target, dtype=np.float64), 0.0)
    loss = float((err ** 2).sum() / count)
    du = 2.0 * err / count
    g = np.zeros(n)
    g[head_w:head_b] = np.tensordot(du, a["d0"], axes=([0, 1, 2, 3], [0, 1, 2, 3]))
    g[head_b] = du.sum()
    e = emulate_bf16

    def back(i, d, act, x, need=True):
        gW, gb, din = conv_backward(d, act, x, p, convs[i], e, need)
        c = convs[i]
        g[c["woff"]:c["woff"] + c["cout"] * c["kp"]] = gW.ravel()
        g[c["boff"]:c["boff"] + c["cout"]] = gb
        return din

    dd0 = du[..., None] * p[head_w:head_b]
    dh0 = back(13, dd0, a["d0"], a["h0"])
    ds0 = dh0.copy()
    du0 = back(12, dh0, a["r0"], a["u0"])
    dd1 = up_backward(du0)
    dh1 = back(11, dd1, a["d1"], a["h1"])
    ds1 = dh1.copy()
    du1 = back(10, dh1, a["r1"], a["u1"])
    dd2 = up_backward(du1)
    dh2 = back(9, dd2, a["d2"], a["h2"])
    ds2 = dh2.copy()
    du2 = back(8, dh2, a["r2"], a["u2"])
    db3 = up_backward(du2)
    da3 = back(7, db3, a["b3"], a["a3"])
    dp3 = back(6, da3, a["a3"], a["p3"])
    ds2 += pool_backward(dp3)
    da2 = back(5, ds2, a["s2"], a["a2"])
    dp2 = back(4, da2, a["a2"], a["p2"])
    ds1 += pool_backward(dp2)
    da1 = back(3, ds1, a["s1"], a["a1"])
    dp1 = back(2, da1, a["a1"], a["p1"])
    ds0 += pool_backward(dp1)
    da0 = back(1, ds0, a["s0"], a["a0"])
    back(0, da0, a["a0"], a["x0"], need=False)
    return loss, g
