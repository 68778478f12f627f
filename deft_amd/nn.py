"""Warm-start network: binding over include/deft_amd_nn.h, training-data
generation with the library's own minimizer, and the training loop.

The network (a periodic 3-D U-Net, BF16 MFMA convolutions in libdeft_amd.so)
predicts u = ln(n / n_bulk) of the converged density from exp(-Vext/kT) and
the bulk packing fraction.  `WarmStartNet.warm_start(fluid, eta)` writes
Veff = -kT (ln n_bulk + u) into a plan, after which the ordinary minimizer
(deft's new-API Polak-Ribiere CG, src/new/Minimize.cpp) runs to convergence;
the network only changes where the minimization starts.

Device memory comes from the library itself (`DeviceArray`), so the network
path loads one HIP runtime only: libdeft_amd.so's.
"""

import ctypes
import math
import os
import time

import numpy as np

from . import (FFT_FP32, DeftError, Minimize, WhiteBearFluidVeff, _check, find_chemical_potential, lib)

_HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(_HERE, "models", "unet_warmstart.npz")
DEFAULT_CHANNELS = (16, 32, 48, 64)
NUM_CONVS = 14


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class DeviceArray:
    """A float32 array in device memory (deft_buffer_create) with a fixed shape."""

    def __init__(self, shape):
        self.shape = tuple(int(d) for d in shape)
        self.size = int(np.prod(self.shape))
        self.nbytes = 4 * self.size
        h = ctypes.c_void_p()
        _check(lib().deft_buffer_create(max(self.nbytes, 4), ctypes.byref(h)))
        self._p = h

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = cls(a.shape)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self._p

    def upload(self, a, offset_elems=0):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if offset_elems < 0 or offset_elems + a.size > self.size:
            raise DeftError("upload out of range")
        _check(lib().deft_buffer_upload(self._p, 4 * offset_elems, a.ctypes.data_as(ctypes.c_void_p), a.nbytes))

    def numpy(self):
        out = np.empty(self.shape, dtype=np.float32)
        _check(lib().deft_buffer_download(self._p, 0, out.ctypes.data_as(ctypes.c_void_p), self.nbytes))
        return out

    def close(self):
        if getattr(self, "_p", None):
            lib().deft_buffer_destroy(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch(boltz, *others):
    """(B, M) of a (B, M, M, M) batch; every other array must have the same shape."""
    if len(boltz.shape) != 4 or not (boltz.shape[1] == boltz.shape[2] == boltz.shape[3]):
        raise DeftError("expected a (B, M, M, M) array")
    for t in others:
        if t is not None and tuple(t.shape) != tuple(boltz.shape):
            raise DeftError("target and mask must have the shape of boltz")
    return boltz.shape[0], boltz.shape[1]


def _etas(eta, B):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(eta, dtype=np.float32), (B,)))


class WarmStartNet:
    """Owns a deft_net (parameters, Adam moments and activation buffers on the device)."""

    def __init__(self, channels=DEFAULT_CHANNELS, seed=0):
        self.channels = tuple(int(c) for c in channels)
        h = ctypes.c_void_p()
        ch = (ctypes.c_int * 4)(*self.channels)
        _check(lib().deft_net_create(ch, int(seed), ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().deft_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_conv_kernel(self, tiled):
        """True: LDS-tiled convolutions where the grid edge allows (default); False: untiled."""
        _check(lib().deft_net_set_conv_kernel(self._h, 1 if tiled else 0))

    @property
    def num_params(self):
        return int(lib().deft_net_num_params(self._h))

    def layout(self, conv):
        """(cin, cout, kp, woff, boff) of convolution `conv` (0..13, -1 = head)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        w, o = ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().deft_net_layout(self._h, int(conv), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                     ctypes.byref(w), ctypes.byref(o)))
        return a.value, b.value, c.value, w.value, o.value

    @property
    def params(self):
        out = np.empty(self.num_params, dtype=np.float32)
        _check(lib().deft_net_get_params(self._h, _fp(out)))
        return out

    @params.setter
    def params(self, p):
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.num_params:
            raise DeftError(f"expected {self.num_params} parameters, got {a.size}")
        _check(lib().deft_net_set_params(self._h, _fp(a)))

    def save(self, path=WEIGHTS_PATH, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params, channels=np.array(self.channels, dtype=np.int32),
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=WEIGHTS_PATH):
        with np.load(path, allow_pickle=False) as z:
            net = cls(tuple(int(c) for c in z["channels"]))
            net.params = z["params"]
        return net

    # -- device calls on DeviceArray batches of shape (B, M, M, M) ---------------

    def forward(self, boltz, eta, out=None):
        """u for a DeviceArray batch of Boltzmann factors; returns a DeviceArray."""
        B, M = _batch(boltz)
        u = out if out is not None else DeviceArray(boltz.shape)
        _check(lib().deft_net_forward(self._h, M, B, boltz.ptr, _fp(_etas(eta, B)), u.ptr))
        return u

    def loss_grad(self, boltz, eta, target, mask=None):
        B, M = _batch(boltz, target, mask)
        loss = ctypes.c_float()
        g = np.empty(self.num_params, dtype=np.float32)
        _check(lib().deft_net_loss_grad(self._h, M, B, boltz.ptr, _fp(_etas(eta, B)), target.ptr,
                                        mask.ptr if mask is not None else None, ctypes.byref(loss), _fp(g)))
        return loss.value, g

    def train_step(self, boltz, eta, target, lr, mask=None):
        B, M = _batch(boltz, target, mask)
        loss = ctypes.c_float()
        _check(lib().deft_net_train_step(self._h, M, B, boltz.ptr, _fp(_etas(eta, B)), target.ptr,
                                         mask.ptr if mask is not None else None, float(lr), ctypes.byref(loss)))
        return loss.value

    def warm_start(self, fluid, eta, factor=1):
        """Sets fluid's Veff from the network's prediction for its Vext; returns the blend
        factor lambda (1: prediction used as is; < 1: pulled toward the bulk start
        because the prediction's energy was not finite).  factor > 1 runs the network on
        a grid `factor` times coarser (for grids finer than the training spacing)."""
        _check(lib().deft_plan_warm_start_scaled(fluid.handle, self._h, float(eta), int(factor)))
        lam = ctypes.c_double()
        _check(lib().deft_net_last_blend(self._h, ctypes.byref(lam)))
        return lam.value


def warm_start_batch(net, fluids, etas, factor=1):
    """WarmStartNet.warm_start for several fluids of one cubic grid edge with a single
    inference over all of them (deft_plans_warm_start); returns the blend factors."""
    n = len(fluids)
    arr = (ctypes.c_void_p * n)(*[f.handle for f in fluids])
    e = (ctypes.c_double * n)(*[float(x) for x in etas])
    lam = (ctypes.c_double * n)()
    _check(lib().deft_plans_warm_start(arr, e, n, net._h, int(factor), lam))
    return [lam[i] for i in range(n)]


def mfma_probe(a, b):
    """c = a @ b.T for a 16x32 and b 16x32 through the convolution kernel's MFMA path."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    if a.shape != (16, 32) or b.shape != (16, 32):
        raise DeftError("mfma_probe takes two 16x32 matrices")
    c = np.empty((16, 16), dtype=np.float32)
    _check(lib().deft_net_mfma_probe(_fp(a), _fp(b), _fp(c)))
    return c


def layer_probe(op, x, channels=None, w=None, bias=None, relu=False, tiled=True, x2=None):
    """Runs one layer kernel on host arrays (deft_net_layer_probe).  op: "conv" (x of
    shape (B, M, M, M, cin), w (channels, kp), output channels `channels`), "dgrad"
    (x = d(pre-activation) with the forward's cout channels, w the forward weights,
    `channels` = the forward's cin), "wgrad" (x = input, x2 = d(pre-activation)),
    "pool" or "up" (x of shape (B, M/2, M/2, M/2, C))."""
    codes = {"conv": 0, "dgrad": 1, "wgrad": 2, "pool": 3, "up": 4}
    x = np.ascontiguousarray(x, dtype=np.float32)
    B, C = x.shape[0], x.shape[-1]
    M = x.shape[1] * (2 if op == "up" else 1)
    if op == "conv":
        cin, cout, shape = C, int(channels), (B, M, M, M, int(channels))
    elif op == "dgrad":
        cin, cout, shape = int(channels), C, (B, M, M, M, int(channels))
    elif op == "wgrad":
        cin, cout = C, int(x2.shape[-1])
        shape = (cout, (27 * cin + 31) // 32 * 32)
    elif op == "pool":
        cin, cout, shape = C, C, (B, M // 2, M // 2, M // 2, C)
    else:
        cin, cout, shape = C, C, (B, M, M, M, C)
    out = np.empty(shape, dtype=np.float32)
    keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float32) for a in (x2, w, bias)]
    _check(lib().deft_net_layer_probe(codes[op], M, B, cin, cout, _fp(x), *(None if a is None else _fp(a) for a in keep),
                                      1 if relu else 0, 1 if tiled else 0, _fp(out)))
    return out


# ---------------------------------------------------------------------------
# training data: random external potentials solved by the library's minimizer
# ---------------------------------------------------------------------------

def _periodic_offsets(M, L, c):
    x = np.arange(M) * (L / M) - c
    return x - L * np.round(x / L)


def random_vext(M, L, rng):
    """A random external potential on an M^3 periodic cell of edge L: hard spheres
    (Vext = +inf where a fluid sphere centre would overlap them), hard slabs and
    soft Gaussian bumps or wells, in random combinations."""
    V = np.zeros((M, M, M))
    kind = int(rng.integers(0, 4))
    for _attempt in range(20):
        V[...] = 0.0
        if kind in (0, 3):
            for _ in range(int(rng.integers(1, 4))):
                c = rng.uniform(0, L, 3)
                rad = rng.uniform(0.3, 3.0)
                dx, dy, dz = (_periodic_offsets(M, L, ci) for ci in c)
                r2 = dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2
                V[r2 < rad * rad] = np.inf
        if kind in (1, 3):
            axis = int(rng.integers(0, 3))
            d = np.abs(_periodic_offsets(M, L, rng.uniform(0, L)))
            t = rng.uniform(0.2, 2.0)
            shape = [1, 1, 1]
            shape[axis] = M
            V = np.where(np.broadcast_to((d < t / 2).reshape(shape), V.shape), np.inf, V)
        if kind == 2 or (kind == 3 and rng.random() < 0.5):
            for _ in range(int(rng.integers(1, 4))):
                c = rng.uniform(0, L, 3)
                s = rng.uniform(0.3, 1.5)
                A = rng.uniform(-0.7, 3.0)
                dx, dy, dz = (_periodic_offsets(M, L, ci) for ci in c)
                r2 = dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2
                V = V + A * np.exp(-r2 / (2 * s * s))
        if np.isfinite(V).mean() > 0.3:
            return V
    return np.zeros((M, M, M))


def training_vext(M, L, rng):
    """Training potentials: half from random_vext, half from the C3 generator
    (8 Gaussians and a planar hard wall, deft_amd.workloads)."""
    from . import workloads
    if rng.random() < 0.5:
        return random_vext(M, L, rng)
    return workloads.c3_cubic_vext(M, L, rng)


def solve(fluid, vext, eta, R=1.0, kT=1.0, relative_precision=1e-10, warm=None):
    """Converges the White-Bear fluid in `vext` at bulk packing fraction eta; returns
    (density, minimizer stats).  `warm` = a WarmStartNet to start from its prediction."""
    nb = eta / (4 * math.pi / 3 * R ** 3)
    fluid.mu = find_chemical_potential(nb, R, kT)
    fluid.Vext = vext
    fluid.fft_precision = FFT_FP32
    if warm is None:
        fluid.fill_Veff(-kT * math.log(nb))
    else:
        warm.warm_start(fluid, eta)
    m = Minimize(fluid)
    m.precondition(True)
    m.set_relative_precision(relative_precision)
    m.set_mixed_precision(1e-6, FFT_FP32)
    m.run()
    s = m.stats()
    s["energy"] = m.energy()
    s["fp32_iterations"], s["mixed_iterations"], _ = m.phase_iterations()
    m.close()
    return fluid.density(), s


def make_dataset(count, M=64, L=6.4, seed=0, eta_range=(0.1, 0.45), R=1.0, kT=1.0, log=None):
    """`count` solved samples as float32 arrays: boltz = exp(-Vext/kT), target = ln(n/n_b)
    (0 where Vext = inf), eta.  Solved with the library's minimizer at relative precision 1e-10."""
    rng = np.random.default_rng(seed)
    fluid = WhiteBearFluidVeff((M, M, M), np.diag([L] * 3), R=R, kT=kT, mu=0.0, fft_precision=FFT_FP32)
    boltz = np.empty((count, M, M, M), dtype=np.float32)
    target = np.zeros((count, M, M, M), dtype=np.float32)
    etas = np.empty(count, dtype=np.float32)
    t0 = time.perf_counter()
    i = 0
    while i < count:
        eta = float(rng.uniform(*eta_range))
        V = training_vext(M, L, rng)
        n, s = solve(fluid, V, eta, R, kT)
        if s["stop_reason"] != "converged" or not np.all(np.isfinite(n)):
            continue
        nb = eta / (4 * math.pi / 3 * R ** 3)
        ok = np.isfinite(V)
        b = np.where(ok, np.exp(-np.where(ok, V, 0.0) / kT), 0.0)
        boltz[i] = b
        target[i] = np.where(ok & (n > 0), np.log(np.where(n > 0, n, 1.0) / nb), 0.0)
        etas[i] = eta
        i += 1
        if log and (i % 100 == 0 or i == count):
            log(f"dataset {i}/{count} samples, {time.perf_counter() - t0:.1f} s")
    fluid.close()
    return {"boltz": boltz, "target": target, "eta": etas, "M": M, "L": L}


def random_transforms(rng, batch, M):
    """Per-grid cubic symmetries and periodic shifts for deft_net_make_batch (9 ints each)."""
    t = np.empty((batch, 9), dtype=np.int32)
    for b in range(batch):
        t[b, :3] = rng.permutation(3)
        t[b, 3:6] = rng.integers(0, 2, 3)
        t[b, 6:] = rng.integers(0, M, 3)
    return t


def make_batch(data_dev, count, M, samples, transforms, out):
    s = np.ascontiguousarray(samples, dtype=np.int32)
    t = np.ascontiguousarray(transforms, dtype=np.int32)
    _check(lib().deft_net_make_batch(data_dev.ptr, int(count), int(M), len(s),
                                     s.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     t.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), out.ptr))


def train(net, data, steps, batch=4, lr=2e-3, lr_final=1e-4, seed=0, log=None, log_every=100):
    """Adam on the masked MSE of u over the dataset (random batches under random cubic
    symmetries and periodic shifts, cosine learning-rate decay).  Returns the loss history."""
    n, M = data["boltz"].shape[0], data["boltz"].shape[1]
    boltz = DeviceArray.from_numpy(data["boltz"])
    target = DeviceArray.from_numpy(data["target"])
    b = DeviceArray((batch, M, M, M))
    t = DeviceArray((batch, M, M, M))
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for step in range(steps):
        idx = rng.integers(0, n, batch)
        xf = random_transforms(rng, batch, M)
        make_batch(boltz, n, M, idx, xf, b)
        make_batch(target, n, M, idx, xf, t)
        cur = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * step / max(steps - 1, 1)))
        hist.append(net.train_step(b, data["eta"][idx], t, cur))
        if log and ((step + 1) % log_every == 0 or step + 1 == steps):
            log(f"step {step + 1}/{steps} loss {np.mean(hist[-log_every:]):.5f} lr {cur:.2e} "
                f"{time.perf_counter() - t0:.1f} s")
    for a in (boltz, target, b, t):
        a.close()
    return hist


def evaluate(net, data):
    """Masked RMS error of u per sample of a dataset."""
    out = []
    for i in range(data["boltz"].shape[0]):
        b = DeviceArray.from_numpy(data["boltz"][i:i + 1])
        u = net.forward(b, data["eta"][i:i + 1]).numpy()[0]
        m = data["boltz"][i] > 0
        out.append(float(np.sqrt(np.mean((u[m] - data["target"][i][m]) ** 2))))
    return out
