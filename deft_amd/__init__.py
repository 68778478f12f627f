"""deft_amd: MI355X-native White-Bear fundamental-measure functional and CG minimizer.

Python binding over the C ABI in ``include/deft_amd.h`` (libdeft_amd.so, built
in-tree by ``python -m deft_amd.build``).  The classes mirror deft's new-API
generated functional and minimizer:

* :class:`WhiteBearFluidVeff` - the generated ``WhiteBearFluidVeff``
  (deft src/haskell/create_generators.py:21-22, src/haskell/NewCode.hs:120-242):
  ``energy()``, ``grad()``, ``energy_grad_and_precond()``, ``Veff``, ``Vext``,
  ``mu``, ``R``, ``kT``;
* :class:`Minimize` - deft's ``Minimize`` (src/new/Minimize.h:13-199):
  ``improve_energy()``, ``set_precision()``, ``set_relative_precision()``,
  ``set_maxiter()``, ``set_miniter()``, ``precondition()``, ``energy()``,
  ``get_iteration_count()``;
* :func:`fft` / :func:`ifft` - the FFT seam (src/Grid.cpp:470-478,
  src/ReciprocalGrid.cpp:52-65);
* :func:`find_chemical_potential`, :func:`pressure` - the bulk path
  (src/equation-of-state.cpp:164-192).

There is no CPU fallback: without the built library or a HIP device every
entry point raises.
"""

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdeft_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "deft_amd.h")
HEADER_PATHS = [HEADER_PATH, os.path.join(os.path.dirname(_HERE), "include", "deft_amd_nn.h")]

FFT_FP32 = 0
FFT_FP64 = 1
WHITE_BEAR = 0
WHITE_BEAR_MARK_II = 1
WHITE_BEAR_MARK_II_SLOW = 2
WHITE_BEAR_TENSOR = 3
SFMT = 4
SW_LIQUID = 5

PGRAD_NEW = 1
PGRAD_EFFECTIVE_POTENTIAL = 2
PGRAD_STRUCTURE_FACTOR = 3

STOP_REASONS = {
    0: "none", 1: "maxiter", 2: "nan-energy", 3: "infinite-energy", 4: "nan-slope",
    5: "zero-slope", 6: "negative-curvature", 7: "energy-rose", 8: "no-change", 9: "converged",
    10: "not-improved", 11: "gradient-unconfirmed",
}


class DeftError(RuntimeError):
    pass


class EnergyTerms(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("kTphi1", "kTphi2", "kTphi3", "ideal", "int_V_n", "minus_mu_N", "total", "sw")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class MinimizeStats(ctypes.Structure):
    _fields_ = [("iter", ctypes.c_int), ("num_energy_calcs", ctypes.c_int), ("num_grad_calcs", ctypes.c_int),
                ("stop_reason", ctypes.c_int), ("energy", ctypes.c_double), ("delta_energy", ctypes.c_double),
                ("error_estimate", ctypes.c_double), ("step", ctypes.c_double)]


_lib = None
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)

_SIGNATURES = {
    "deft_last_error": ([], ctypes.c_char_p),
    "deft_version": ([], _I),
    "deft_supported_length": ([_I], _I),
    "deft_fused_length": ([_I], _I),
    "deft_device_count": ([_IP], _I),
    "deft_plan_create": ([_I, _I, _I, _DP, _D, _D, _D, _I, ctypes.POINTER(_P)], _I),
    "deft_plan_destroy": ([_P], _I),
    "deft_plan_autotune": ([_P], _I),
    "deft_plan_tuning": ([_P, _I, _IP, _IP, _IP], _I),
    "deft_plan_set_tuning": ([_P, _I, _I, _I, _I], _I),
    "deft_plan_zphi_variant": ([_P, _I, _IP], _I),
    "deft_plan_set_zphi_variant": ([_P, _I, _I], _I),
    "deft_plan_xfwd_variant": ([_P, _I, _IP], _I),
    "deft_plan_set_xfwd_variant": ([_P, _I, _I], _I),
    "deft_plan_radix_schedule": ([_P, _I, _I, _IP, _I, _IP], _I),
    "deft_plan_set_radix_schedule": ([_P, _I, _I, _IP, _I], _I),
    "deft_plan_set_fft_precision": ([_P, _I], _I),
    "deft_plan_set_mu": ([_P, _D], _I),
    "deft_plan_bulk_density": ([_P, _DP], _I),
    "deft_plan_set_functional": ([_P, _I], _I),
    "deft_plan_set_sfmt": ([_P, _D, _D, _D], _I),
    "deft_sfmt_homogeneous": ([_D, _D, _D, _D, _D, _D, _DP, _DP], _I),
    "deft_plan_set_sw": ([_P, _D, _D], _I),
    "deft_sw_homogeneous": ([_D, _D, _D, _D, _D, _D, _DP, _DP], _I),
    "deft_contact_values": ([_P, _DP], _I),
    "deft_copy_bandwidth": ([ctypes.c_longlong, _I, _DP], _I),
    "deft_bulk_state": ([_I, _D, _D, _D, _DP, _DP, _DP], _I),
    "deft_plan_set_vext": ([_P, _DP], _I),
    "deft_plan_set_veff": ([_P, _DP], _I),
    "deft_plan_get_veff": ([_P, _DP], _I),
    "deft_plan_fill_veff": ([_P, _D], _I),
    "deft_plan_get_density": ([_P, _DP], _I),
    "deft_plan_veff_device_ptr": ([_P, ctypes.POINTER(_DP)], _I),
    "deft_plan_sizes": ([_P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)], _I),
    "deft_functional_energy": ([_P, _DP, _DP, ctypes.POINTER(EnergyTerms)], _I),
    "deft_functional_grad": ([_P, _DP, _DP, _DP, _DP], _I),
    "deft_functional_grad_device": ([_P, _DP, _DP, _DP, _DP], _I),
    "deft_functional_grad_kind": ([_P, _DP, _DP, _DP, _DP, _I], _I),
    "deft_functional_grad_device_kind": ([_P, _DP, _DP, _DP, _DP, _I], _I),
    "deft_functional_grad_noalloc": ([_I, _I, _I, _DP, _D, _D, _D, _I, _I, _P, _P, _P, _P, _P, _P], _I),
    "deft_minimize_set_gradient_check": ([_P, _D], _I),
    "deft_minimize_gradient_check_get": ([_P, _DP, _IP], _I),
    "deft_cg_create": ([_P, _I, _D, ctypes.POINTER(_P)], _I),
    "deft_cg_destroy": ([_P], _I),
    "deft_cg_set_precision": ([_P, _D], _I),
    "deft_cg_set_maxiter": ([_P, _I], _I),
    "deft_cg_set_fft_precision": ([_P, _I], _I),
    "deft_cg_improve_energy": ([_P, _I, _IP], _I),
    "deft_cg_energy": ([_P, _DP], _I),
    "deft_cg_grad": ([_P, _DP], _I),
    "deft_cg_pgrad": ([_P, _DP], _I),
    "deft_cg_minimize": ([_P], _I),
    "deft_cg_invalidate": ([_P], _I),
    "deft_cg_stats_get": ([_P, ctypes.POINTER(MinimizeStats)], _I),
    "deft_weighted_densities": ([_P, _DP], _I),
    "deft_fft": ([_P, _DP, _DP], _I),
    "deft_ifft": ([_P, _DP, _DP], _I),
    "deft_bulk_chemical_potential": ([_D, _D, _D, _DP], _I),
    "deft_bulk_pressure": ([_D, _D, _D, _DP], _I),
    "deft_bulk_free_energy_density": ([_D, _D, _D, _DP], _I),
    "deft_minimize_create": ([_P, ctypes.POINTER(_P)], _I),
    "deft_minimize_destroy": ([_P], _I),
    "deft_minimize_set_precision": ([_P, _D], _I),
    "deft_minimize_set_relative_precision": ([_P, _D], _I),
    "deft_minimize_set_maxiter": ([_P, _I], _I),
    "deft_minimize_set_miniter": ([_P, _I], _I),
    "deft_minimize_precondition": ([_P, _I], _I),
    "deft_minimize_set_preconditioner": ([_P, _I], _I),
    "deft_minimize_improve_energy": ([_P, _I, _IP], _I),
    "deft_minimize_run": ([_P, _I, _IP], _I),
    "deft_minimize_energy": ([_P, _DP], _I),
    "deft_minimize_stats_get": ([_P, ctypes.POINTER(MinimizeStats)], _I),
    "deft_minimize_invalidate": ([_P], _I),
    "deft_minimize_set_mixed_precision": ([_P, _D, _I], _I),
    "deft_minimize_set_fused_line_search": ([_P, _I], _I),
    "deft_minimize_fp32_iterations": ([_P, _IP], _I),
    "deft_minimize_phase_iterations": ([_P, _IP, _IP, _IP], _I),
    "deft_plan_set_profiling": ([_P, _I], _I),
    "deft_plan_profile_reset": ([_P], _I),
    "deft_kernel_count": ([], _I),
    "deft_plan_profile_get": ([_P, _I, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_longlong), _DP, _DP], _I),
    "deft_batch_create": ([ctypes.POINTER(_P), _I, ctypes.POINTER(_P)], _I),
    "deft_batch_minimize": ([_P, ctypes.POINTER(_P), _I], _I),
    "deft_batch_set_graphs": ([_P, _I], _I),
    "deft_batch_stats": ([_P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                          ctypes.POINTER(ctypes.c_longlong)], _I),
    "deft_batch_destroy": ([_P], _I),
    # warm-start network (include/deft_amd_nn.h); device buffers are passed as integers
    "deft_net_create": ([_IP, ctypes.c_ulonglong, ctypes.POINTER(_P)], _I),
    "deft_net_destroy": ([_P], _I),
    "deft_net_num_params": ([_P], ctypes.c_longlong),
    "deft_net_set_conv_kernel": ([_P, _I], _I),
    "deft_net_layout": ([_P, _I, _IP, _IP, _IP, ctypes.POINTER(ctypes.c_longlong),
                         ctypes.POINTER(ctypes.c_longlong)], _I),
    "deft_net_get_params": ([_P, _FP], _I),
    "deft_net_set_params": ([_P, _FP], _I),
    "deft_net_forward": ([_P, _I, _I, _P, _FP, _P], _I),
    "deft_net_loss_grad": ([_P, _I, _I, _P, _FP, _P, _P, _FP, _FP], _I),
    "deft_net_train_step": ([_P, _I, _I, _P, _FP, _P, _P, ctypes.c_float, _FP], _I),
    "deft_plan_warm_start": ([_P, _P, _D], _I),
    "deft_plan_warm_start_scaled": ([_P, _P, _D, _I], _I),
    "deft_net_last_blend": ([_P, _DP], _I),
    "deft_plan_engine": ([_P, _IP, _DP, _DP], _I),
    "deft_plan_set_engine": ([_P, _I], _I),
    "deft_plans_warm_start": ([ctypes.POINTER(_P), _DP, _I, _P, _I, _DP], _I),
    "deft_buffer_create": ([ctypes.c_ulonglong, ctypes.POINTER(_P)], _I),
    "deft_buffer_destroy": ([_P], _I),
    "deft_buffer_upload": ([_P, ctypes.c_ulonglong, _P, ctypes.c_ulonglong], _I),
    "deft_buffer_download": ([_P, ctypes.c_ulonglong, _P, ctypes.c_ulonglong], _I),
    "deft_net_make_batch": ([_P, ctypes.c_longlong, _I, _I, _IP, _IP, _P], _I),
    "deft_net_mfma_probe": ([_FP, _FP, _FP], _I),
    "deft_net_layer_probe": ([_I, _I, _I, _I, _I, _FP, _FP, _FP, _FP, _I, _I, _FP], _I),
}


def lib():
    """Loads libdeft_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DeftError(f"{LIB_PATH} is missing: run `python -m deft_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().deft_last_error()
        raise DeftError(f"deft_amd error {rc}: {msg.decode() if msg else ''}")


def _dptr(a):
    return a.ctypes.data_as(_DP)


def device_count():
    c = ctypes.c_int(0)
    _check(lib().deft_device_count(ctypes.byref(c)))
    return c.value


def supported_length(n):
    return bool(lib().deft_supported_length(int(n)))


def fused_length(n):
    """True when the fused per-edge kernels serve this edge (other supported
    edges run the generic line passes)."""
    return bool(lib().deft_fused_length(int(n)))


# ---------------------------------------------------------------------------
# bulk path
# ---------------------------------------------------------------------------

def copy_bandwidth(nbytes=1 << 30, reps=20):
    """Attainable HBM bandwidth (GB/s) of a streaming device copy."""
    out = ctypes.c_double()
    _check(lib().deft_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(out)))
    return out.value


def bulk_state(n, R=1.0, kT=1.0, functional=WHITE_BEAR):
    """(f, mu, p) of ideal gas + the hard-sphere functional at bulk density n."""
    f, mu, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _check(lib().deft_bulk_state(int(functional), n, R, kT, ctypes.byref(f), ctypes.byref(mu), ctypes.byref(p)))
    return f.value, mu.value, p.value


def find_chemical_potential(n, R=1.0, kT=1.0, functional=WHITE_BEAR):
    """mu of ideal gas + White-Bear at bulk density n (src/equation-of-state.cpp:189-192)."""
    if functional != WHITE_BEAR:
        return bulk_state(n, R, kT, functional)[1]
    out = ctypes.c_double()
    _check(lib().deft_bulk_chemical_potential(n, R, kT, ctypes.byref(out)))
    return out.value


def pressure(n, R=1.0, kT=1.0, functional=WHITE_BEAR):
    """Bulk pressure (src/equation-of-state.cpp:164-170)."""
    if functional != WHITE_BEAR:
        return bulk_state(n, R, kT, functional)[2]
    out = ctypes.c_double()
    _check(lib().deft_bulk_pressure(n, R, kT, ctypes.byref(out)))
    return out.value


def bulk_free_energy_density(n, R=1.0, kT=1.0):
    out = ctypes.c_double()
    _check(lib().deft_bulk_free_energy_density(n, R, kT, ctypes.byref(out)))
    return out.value


class HomogeneousSFMTFluid:
    """The generated HomogeneousSFMTFluid (src/haskell/SFMT.hs:39-41): a uniform
    soft-FMT fluid with settable sigma, epsilon, Xi, kT, n and mu; ``energy()`` is
    the free energy per volume kT Phi + kT n (ln(n/nQ) - 1) - mu n and
    ``d_by_dn()`` its derivative (drivers set mu = d_by_dn() at mu = 0,
    papers/fuzzy-fmt/figs/new-radial-wca.cpp:121-123)."""

    def __init__(self, sigma=1.0, epsilon=1.0, Xi=0.1, kT=1.0, n=0.0, mu=0.0):
        self.sigma, self.epsilon, self.Xi, self.kT, self.n, self.mu = sigma, epsilon, Xi, kT, n, mu

    def _eval(self):
        e, d = ctypes.c_double(), ctypes.c_double()
        _check(lib().deft_sfmt_homogeneous(float(self.sigma), float(self.epsilon), float(self.Xi), float(self.kT),
                                           float(self.n), float(self.mu), ctypes.byref(e), ctypes.byref(d)))
        return e.value, d.value

    def energy(self):
        return self._eval()[0]

    def d_by_dn(self):
        return self._eval()[1]

    def pressure(self):
        """n d_by_dn - energy (papers/fuzzy-fmt/figs/homogeneous.cpp:57), at the set mu."""
        e, d = self._eval()
        return self.n * d - e


class HomogeneousSW_liquid:
    """The generated HomogeneousSW_liquid (src/haskell/SW_liquid.hs:45-47): the
    uniform square-well liquid with settable R, epsilon, lambda, kT, n and mu;
    ``energy()`` is the free energy per volume, ``d_by_dn()`` its derivative
    (papers/square-well-fluid/figs/new-radial-sw.cpp:113-125 sets mu = d_by_dn()
    at mu = 0)."""

    def __init__(self, R=1.0, epsilon=1.0, lambda_=1.3, kT=1.0, n=0.0, mu=0.0):
        self.R, self.epsilon, self.lambda_, self.kT, self.n, self.mu = R, epsilon, lambda_, kT, n, mu

    def _eval(self):
        e, d = ctypes.c_double(), ctypes.c_double()
        _check(lib().deft_sw_homogeneous(float(self.R), float(self.epsilon), float(self.lambda_), float(self.kT),
                                         float(self.n), float(self.mu), ctypes.byref(e), ctypes.byref(d)))
        return e.value, d.value

    def energy(self):
        return self._eval()[0]

    def d_by_dn(self):
        return self._eval()[1]

    def pressure(self):
        e, d = self._eval()
        return self.n * d - e


# ---------------------------------------------------------------------------
# functional
# ---------------------------------------------------------------------------

class WhiteBearFluidVeff:
    """White-Bear fluid of Veff on one grid, resident on the GPU.

    ``shape`` = (Nx, Ny, Nz), each edge 1..4096 with no prime factor above 512
    (``supported_length``; edges with ``fused_length`` run the fused passes);
    ``cell`` = 3x3 lattice vectors as rows (or three edge lengths of an
    orthorhombic cell).  ``Vext`` is FP64 on the host, kept in FP64 for the FP64
    passes and rounded to FP32 for the FP32 ones; +inf marks hard walls, NaN
    constrained points.
    """

    def __init__(self, shape, cell, R=1.0, kT=1.0, mu=0.0, fft_precision=FFT_FP32, functional=WHITE_BEAR):
        self.shape = tuple(int(s) for s in shape)
        cell = np.asarray(cell, dtype=np.float64)
        if cell.shape == (3,):
            cell = np.diag(cell)
        self.cell = np.ascontiguousarray(cell.reshape(3, 3))
        self._R, self._kT, self._mu = float(R), float(kT), float(mu)
        self._prec = int(fft_precision)
        h = ctypes.c_void_p()
        _check(lib().deft_plan_create(self.shape[0], self.shape[1], self.shape[2], _dptr(self.cell),
                                      self._R, self._kT, self._mu, self._prec, ctypes.byref(h)))
        self._h = h
        if functional != WHITE_BEAR:
            _check(lib().deft_plan_set_functional(self._h, int(functional)))
        self._functional = int(functional)
        self.N = int(np.prod(self.shape))
        self.kshape = (self.shape[0], self.shape[1], self.shape[2] // 2 + 1)
        self.volume = abs(float(np.linalg.det(self.cell)))
        self.dV = self.volume / self.N
        self._vext = np.zeros(self.shape)
        self.last_terms = {}

    @classmethod
    def from_spacing(cls, ax, ay, az, dx, **kw):
        """The generated (ax, ay, az, dx) constructor: N = ceil(a/dx/2)*2 (src/haskell/NewCode.hs:160-168)."""
        shape = [int(math.ceil(a / dx / 2)) * 2 for a in (ax, ay, az)]
        return cls(shape, np.diag([ax, ay, az]), **kw)

    def close(self):
        if getattr(self, "_h", None):
            lib().deft_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # parameters -----------------------------------------------------------
    @property
    def R(self):
        return self._R

    @property
    def kT(self):
        return self._kT

    @property
    def mu(self):
        return self._mu

    @mu.setter
    def mu(self, v):
        _check(lib().deft_plan_set_mu(self._h, float(v)))
        self._mu = float(v)

    @property
    def fft_precision(self):
        return self._prec

    @fft_precision.setter
    def fft_precision(self, v):
        _check(lib().deft_plan_set_fft_precision(self._h, int(v)))
        self._prec = int(v)

    @property
    def Vext(self):
        return self._vext.copy()

    @Vext.setter
    def Vext(self, v):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), self.shape))
        _check(lib().deft_plan_set_vext(self._h, _dptr(a)))
        self._vext = a.copy()

    @property
    def Veff(self):
        out = np.empty(self.shape)
        _check(lib().deft_plan_get_veff(self._h, _dptr(out)))
        return out

    @Veff.setter
    def Veff(self, v):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), self.shape))
        _check(lib().deft_plan_set_veff(self._h, _dptr(a)))

    @property
    def functional(self):
        """WHITE_BEAR (HardSpheresWBnotensor), WHITE_BEAR_MARK_II (WhiteBearMarkII) or
        WHITE_BEAR_MARK_II_SLOW (HardSpheresWBm2slow: Phi3 over 36 pi, not 24 pi) or
        WHITE_BEAR_TENSOR (TensorWhiteBear)."""
        return self._functional

    @functional.setter
    def functional(self, v):
        _check(lib().deft_plan_set_functional(self._h, int(v)))
        self._functional = int(v)

    def fill_Veff(self, value):
        """Veff = value everywhere, written on the device."""
        _check(lib().deft_plan_fill_veff(self._h, float(value)))

    def density(self):
        out = np.empty(self.shape)
        _check(lib().deft_plan_get_density(self._h, _dptr(out)))
        return out

    # functional seam ------------------------------------------------------
    def energy(self, Veff=None):
        if Veff is not None:
            self.Veff = Veff
        e = ctypes.c_double()
        t = EnergyTerms()
        _check(lib().deft_functional_energy(self._h, None, ctypes.byref(e), ctypes.byref(t)))
        self.last_terms = t.as_dict()
        if self._functional != SW_LIQUID:
            del self.last_terms["sw"]  # the White-Bear family's printme has six terms
        return e.value

    def grad(self, Veff=None):
        return self.energy_grad_and_precond(Veff)[1]

    def integralgrad(self, Veff=None):
        """(energy, grad, pgrad) with the old API's pgrad (dF/dn)/(Veff - kT)
        (Functional::integralgrad, src/Functional.h:144-149; src/EffectivePotentialToDensity.cpp:62-71)."""
        if Veff is not None:
            self.Veff = Veff
        g = np.zeros(self.shape)
        pg = np.zeros(self.shape)
        e = ctypes.c_double()
        _check(lib().deft_functional_grad_kind(self._h, None, _dptr(g), _dptr(pg), ctypes.byref(e),
                                               PGRAD_EFFECTIVE_POTENTIAL))
        return e.value, g, pg

    def sk_precond(self, Veff=None):
        """(energy, grad, pgrad) with the structure-factor preconditioned gradient
        (DEFT_PGRAD_STRUCTURE_FACTOR; no counterpart in the reference)."""
        if Veff is not None:
            self.Veff = Veff
        g = np.zeros(self.shape)
        pg = np.zeros(self.shape)
        e = ctypes.c_double()
        _check(lib().deft_functional_grad_kind(self._h, None, _dptr(g), _dptr(pg), ctypes.byref(e),
                                               PGRAD_STRUCTURE_FACTOR))
        return e.value, g, pg

    def bulk_density(self):
        """The bulk density whose chemical potential is this fluid's mu."""
        v = ctypes.c_double()
        _check(lib().deft_plan_bulk_density(self._h, ctypes.byref(v)))
        return v.value

    def energy_grad_and_precond(self, Veff=None):
        if Veff is not None:
            self.Veff = Veff
        g = np.zeros(self.shape)
        pg = np.zeros(self.shape)
        e = ctypes.c_double()
        _check(lib().deft_functional_grad(self._h, None, _dptr(g), _dptr(pg), ctypes.byref(e)))
        return e.value, g, pg

    def autotune(self):
        """Times the pass variants on this grid and keeps the fastest (like FFTW_MEASURE)."""
        _check(lib().deft_plan_autotune(self._h))
        return {p: self.tuning(p) for p in (FFT_FP32, FFT_FP64)}

    def engine(self):
        """("generic" or "fused", fused ms, generic ms): the engine the plan runs and the
        energy+gradient evaluation times (FP32 + FP64) the last autotune measured."""
        g, a, b = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        _check(lib().deft_plan_engine(self._h, ctypes.byref(g), ctypes.byref(a), ctypes.byref(b)))
        return ("generic" if g.value else "fused"), a.value, b.value

    def set_engine(self, generic):
        _check(lib().deft_plan_set_engine(self._h, 1 if generic else 0))

    def tuning(self, fft_precision):
        z, y, x = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().deft_plan_tuning(self._h, int(fft_precision), ctypes.byref(z), ctypes.byref(y), ctypes.byref(x)))
        w, xw = ctypes.c_int(), ctypes.c_int()
        _check(lib().deft_plan_zphi_variant(self._h, int(fft_precision), ctypes.byref(w)))
        _check(lib().deft_plan_xfwd_variant(self._h, int(fft_precision), ctypes.byref(xw)))
        return {"zblock": z.value, "ytile": y.value, "xtile": x.value, "zphi": w.value, "xfwd": xw.value}

    def set_tuning(self, fft_precision, zblock, ytile, xtile, zphi=None, xfwd=None):
        """Explicit pass variants; the Phi pass layout `zphi` and the forward x pass
        layout `xfwd` (0 block-wide FFT stages, 1 per wave) are left as they are when None."""
        _check(lib().deft_plan_set_tuning(self._h, int(fft_precision), int(zblock), int(ytile), int(xtile)))
        if zphi is not None:
            _check(lib().deft_plan_set_zphi_variant(self._h, int(fft_precision), int(zphi)))
        if xfwd is not None:
            _check(lib().deft_plan_set_xfwd_variant(self._h, int(fft_precision), int(xfwd)))

    def radix_schedule(self, fft_precision, axis):
        """The radix order (Stockham stages) of one axis's transforms."""
        buf = (ctypes.c_int * 16)()
        n = ctypes.c_int()
        _check(lib().deft_plan_radix_schedule(self._h, int(fft_precision), int(axis), buf, 16, ctypes.byref(n)))
        return [buf[i] for i in range(n.value)]

    def set_radix_schedule(self, fft_precision, axis, radices):
        arr = (ctypes.c_int * max(1, len(radices)))(*[int(r) for r in radices])
        _check(lib().deft_plan_set_radix_schedule(self._h, int(fft_precision), int(axis), arr, len(radices)))

    def set_profiling(self, on=True):
        _check(lib().deft_plan_set_profiling(self._h, 1 if on else 0))

    def profile_reset(self):
        _check(lib().deft_plan_profile_reset(self._h))

    def profile(self):
        """{kernel: (launches, total_ms, total_bytes)} accumulated while profiling was on."""
        out = {}
        for k in range(lib().deft_kernel_count()):
            name = ctypes.c_char_p()
            n = ctypes.c_longlong()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            _check(lib().deft_plan_profile_get(self._h, k, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms),
                                               ctypes.byref(by)))
            out[name.value.decode()] = (n.value, ms.value, by.value)
        return out

    def contact_values(self):
        """n0, gSigmaS, YuWu_S, nA, gSigmaA, Gross of the resident density (the fields the
        contact-paper drivers write beside it, papers/contact/figs/inner-sphere.cpp:220-233)."""
        out = np.empty((7,) + self.shape)
        _check(lib().deft_contact_values(self._h, _dptr(out)))
        return {k: out[i] for i, k in enumerate(("n0", "gSigmaS", "YuWu_S", "nA", "gSigmaA", "Gross", "gSigmaS_cpp"))}

    def weighted_densities(self):
        nw = {WHITE_BEAR_TENSOR: 11, SFMT: 16}.get(self._functional, 5)
        out = np.empty((nw,) + self.shape)
        _check(lib().deft_weighted_densities(self._h, _dptr(out)))
        return out

    def fft(self, f):
        """deft's fft(gd, f) = r2c * dV (src/Grid.cpp:470-478)."""
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.shape)
        out = np.empty(self.kshape, dtype=np.complex128)
        _check(lib().deft_fft(self._h, _dptr(a), out.ctypes.data_as(_DP)))
        return out

    def ifft(self, F):
        """deft's ifft(gd, F) = c2r / V (src/ReciprocalGrid.cpp:52-65)."""
        a = np.ascontiguousarray(F, dtype=np.complex128).reshape(self.kshape)
        out = np.empty(self.shape)
        _check(lib().deft_ifft(self._h, a.ctypes.data_as(_DP), _dptr(out)))
        return out


class SFMTFluidVeff(WhiteBearFluidVeff):
    """The generated SFMTFluidVeff (src/haskell/SFMT.hs:35-37,
    src/haskell/create_generators.py:25-26): soft FMT of Veff on one grid,
    with the sigma, epsilon and Xi accessors of the generated class (alpha
    follows from sigma, epsilon and kT).  Runs on the generic line passes."""

    def __init__(self, shape, cell, sigma=1.0, epsilon=1.0, Xi=0.1, kT=1.0, mu=0.0, fft_precision=FFT_FP32):
        super().__init__(shape, cell, R=sigma / 2, kT=kT, mu=mu, fft_precision=fft_precision)
        self._sigma, self._epsilon, self._Xi = float(sigma), float(epsilon), float(Xi)
        _check(lib().deft_plan_set_sfmt(self._h, self._sigma, self._epsilon, self._Xi))
        self.functional = SFMT

    def _set(self):
        _check(lib().deft_plan_set_sfmt(self._h, self._sigma, self._epsilon, self._Xi))

    @property
    def sigma(self):
        return self._sigma

    @sigma.setter
    def sigma(self, v):
        self._sigma = float(v)
        self._set()

    @property
    def epsilon(self):
        return self._epsilon

    @epsilon.setter
    def epsilon(self, v):
        self._epsilon = float(v)
        self._set()

    @property
    def Xi(self):
        return self._Xi

    @Xi.setter
    def Xi(self, v):
        self._Xi = float(v)
        self._set()

    @property
    def alpha(self):
        """sigma (2/(1 + sqrt(ln 2 kT/epsilon)))^(1/6) (src/haskell/SFMT.hs:55-56)."""
        return self._sigma * (2.0 / (1.0 + math.sqrt(math.log(2.0) * self._kT / self._epsilon))) ** (1.0 / 6.0)


class SWLiquidVeff(WhiteBearFluidVeff):
    """The generated SW_liquidVeff (src/haskell/SW_liquid.hs:13-27,
    src/haskell/create_generators.py): White-Bear of Veff plus the square-well
    attraction of depth ``epsilon`` and range ``lambda_`` sigma, sigma = 2R.
    Runs on the generic line passes; ``last_terms['sw']`` is the attraction."""

    def __init__(self, shape, cell, R=1.0, epsilon=1.0, lambda_=1.3, kT=1.0, mu=0.0, fft_precision=FFT_FP32):
        super().__init__(shape, cell, R=R, kT=kT, mu=mu, fft_precision=fft_precision)
        self._epsilon, self._lambda = float(epsilon), float(lambda_)
        _check(lib().deft_plan_set_sw(self._h, self._epsilon, self._lambda))
        self.functional = SW_LIQUID

    @property
    def epsilon(self):
        return self._epsilon

    @epsilon.setter
    def epsilon(self, v):
        self._epsilon = float(v)
        _check(lib().deft_plan_set_sw(self._h, self._epsilon, self._lambda))

    @property
    def lambda_(self):
        return self._lambda

    @lambda_.setter
    def lambda_(self, v):
        self._lambda = float(v)
        _check(lib().deft_plan_set_sw(self._h, self._epsilon, self._lambda))


class Minimize:
    """deft's new-API minimizer (src/new/Minimize.h:13-199) on a :class:`WhiteBearFluidVeff`."""

    def __init__(self, f):
        self.f = f
        h = ctypes.c_void_p()
        _check(lib().deft_minimize_create(f.handle, ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().deft_minimize_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, p):
        _check(lib().deft_minimize_set_precision(self._h, float(p)))

    def set_relative_precision(self, p):
        _check(lib().deft_minimize_set_relative_precision(self._h, float(p)))

    def set_maxiter(self, m):
        _check(lib().deft_minimize_set_maxiter(self._h, int(m)))

    def set_miniter(self, m):
        _check(lib().deft_minimize_set_miniter(self._h, int(m)))

    def precondition(self, u):
        _check(lib().deft_minimize_precondition(self._h, 1 if u else 0))

    def set_preconditioner(self, kind):
        """PGRAD_NEW (grad/n, the reference's) or PGRAD_STRUCTURE_FACTOR."""
        _check(lib().deft_minimize_set_preconditioner(self._h, int(kind)))

    def improve_energy(self, verbose=False):
        more = ctypes.c_int()
        _check(lib().deft_minimize_improve_energy(self._h, 1 if verbose else 0, ctypes.byref(more)))
        return bool(more.value)

    def run(self, max_calls=1000000):
        it = ctypes.c_int()
        _check(lib().deft_minimize_run(self._h, int(max_calls), ctypes.byref(it)))
        return it.value

    def energy(self):
        e = ctypes.c_double()
        _check(lib().deft_minimize_energy(self._h, ctypes.byref(e)))
        return e.value

    def stats(self):
        s = MinimizeStats()
        _check(lib().deft_minimize_stats_get(self._h, ctypes.byref(s)))
        d = {n: getattr(s, n) for n, _ in s._fields_}
        d["stop_reason"] = STOP_REASONS.get(d["stop_reason"], str(d["stop_reason"]))
        return d

    def get_iteration_count(self):
        return self.stats()["iter"]

    def set_mixed_precision(self, switch_relative_precision, polish_grad_precision=FFT_FP32):
        """FP32 FFTs until this relative precision (or an FP32 stall), then FP64 line-search
        energies with gradients at ``polish_grad_precision`` (0 disables)."""
        _check(lib().deft_minimize_set_mixed_precision(self._h, float(switch_relative_precision),
                                                       int(polish_grad_precision)))

    def fp32_iterations(self):
        it = ctypes.c_int()
        _check(lib().deft_minimize_fp32_iterations(self._h, ctypes.byref(it)))
        return it.value

    def phase_iterations(self):
        """(phase-0 FP32 iterations, phase-1 mixed iterations, current phase)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().deft_minimize_phase_iterations(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_gradient_check(self, factor):
        """Gradient confirmation of the ladder's final verdict (0 disables)."""
        _check(lib().deft_minimize_set_gradient_check(self._h, float(factor)))

    def gradient_check(self):
        """(Newton error estimate of the last confirmation, restarts it caused)."""
        e, r = ctypes.c_double(), ctypes.c_int()
        _check(lib().deft_minimize_gradient_check_get(self._h, ctypes.byref(e), ctypes.byref(r)))
        return e.value, r.value

    def set_fused_line_search(self, mode):
        """1: energy+gradient at the quadratic step's trial point, 0: energy only,
        -1: on with the mixed-precision ladder (default)."""
        _check(lib().deft_minimize_set_fused_line_search(self._h, int(mode)))

    def invalidate_cache(self):
        _check(lib().deft_minimize_invalidate(self._h))


class ConjugateGradient:
    """deft's old-API minimizer stack (MinimizerInterface, src/Minimizer.h:16-81) on a
    :class:`WhiteBearFluidVeff`: ``ConjugateGradient(f, gd, kT, &x, QuadraticLineMinimizer,
    stepsize)`` (src/ConjugateGradient.cpp:59-108), or with ``preconditioned``
    ``PreconditionedConjugateGradient`` (:115-160), optionally wrapped in ``Precision`` and
    ``MaxIter`` (see :func:`Precision`, :func:`MaxIter`).  x is the fluid's resident Veff."""

    def __init__(self, f, preconditioned=False, stepsize=10.0):
        self.f = f
        h = ctypes.c_void_p()
        _check(lib().deft_cg_create(f.handle, 1 if preconditioned else 0, float(stepsize), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().deft_cg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, p):
        _check(lib().deft_cg_set_precision(self._h, float(p)))

    def set_maxiter(self, n):
        _check(lib().deft_cg_set_maxiter(self._h, int(n)))

    def set_fft_precision(self, p):
        _check(lib().deft_cg_set_fft_precision(self._h, int(p)))

    def improve_energy(self, verbose=False):
        more = ctypes.c_int()
        _check(lib().deft_cg_improve_energy(self._h, 1 if verbose else 0, ctypes.byref(more)))
        return bool(more.value)

    def energy(self):
        e = ctypes.c_double()
        _check(lib().deft_cg_energy(self._h, ctypes.byref(e)))
        return e.value

    def grad(self):
        out = np.empty(self.f.shape)
        _check(lib().deft_cg_grad(self._h, _dptr(out)))
        return out

    def pgrad(self):
        out = np.empty(self.f.shape)
        _check(lib().deft_cg_pgrad(self._h, _dptr(out)))
        return out

    def minimize(self):
        """minimize(f, gd, x): iteration count, step and direction start over."""
        _check(lib().deft_cg_minimize(self._h))

    def invalidate_cache(self):
        _check(lib().deft_cg_invalidate(self._h))

    def stats(self):
        s = MinimizeStats()
        _check(lib().deft_cg_stats_get(self._h, ctypes.byref(s)))
        d = {n: getattr(s, n) for n, _ in s._fields_}
        d["stop_reason"] = STOP_REASONS.get(d["stop_reason"], str(d["stop_reason"]))
        return d


def PreconditionedConjugateGradient(f, stepsize=10.0):
    """PreconditionedConjugateGradient(f, gd, kT, &x, QuadraticLineMinimizer, stepsize)."""
    return ConjugateGradient(f, preconditioned=True, stepsize=stepsize)


def Precision(precision, m):
    """Precision(precision, m) (src/Precision.cpp:23-156) around an old-API minimizer."""
    m.set_precision(precision)
    return m


def MaxIter(maxiter, m):
    """MaxIter(maxiter, m) (src/Minimizer.cpp) around an old-API minimizer."""
    m.set_maxiter(maxiter)
    return m


def functional_grad_noalloc(shape, cell, R, kT, mu, fft_precision, pgrad_kind, veff_dev, vext_dev, outgrad_dev,
                            outpgrad_dev, energy_dev, stream=None):
    """deft_functional_grad_noalloc on device buffers given as integers (addresses); no
    allocation, enqueued on `stream` (an integer hipStream_t or None) without waiting."""
    cell = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3))
    _check(lib().deft_functional_grad_noalloc(int(shape[0]), int(shape[1]), int(shape[2]), _dptr(cell), float(R),
                                              float(kT), float(mu), int(fft_precision), int(pgrad_kind),
                                              veff_dev, vext_dev, outgrad_dev, outpgrad_dev, energy_dev, stream))


class Batch:
    """Several fluids minimized together (deft_batch_*): every evaluation the
    minimizers ask for runs in a lockstep round of batched launches."""

    def __init__(self, fluids, graphs=False):
        self.fluids = list(fluids)
        arr = (_P * len(self.fluids))(*[f.handle for f in self.fluids])
        h = ctypes.c_void_p()
        _check(lib().deft_batch_create(arr, len(self.fluids), ctypes.byref(h)))
        self._h = h
        if graphs:
            _check(lib().deft_batch_set_graphs(self._h, 1))

    def minimize(self, minimizers):
        """Runs each minimizer (one per fluid of the batch) to its own stop."""
        arr = (_P * len(minimizers))(*[m._h for m in minimizers])
        _check(lib().deft_batch_minimize(self._h, arr, len(minimizers)))

    def stats(self):
        r, l, q = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().deft_batch_stats(self._h, ctypes.byref(r), ctypes.byref(l), ctypes.byref(q)))
        return {"rounds": r.value, "launches": l.value, "requests": q.value}

    def close(self):
        if getattr(self, "_h", None):
            _check(lib().deft_batch_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def declared_symbols(headers=None):
    """Function names declared in include/deft_amd.h and include/deft_amd_nn.h."""
    import re
    names = set()
    for h in headers or HEADER_PATHS:
        names |= set(re.findall(r"\b(deft_[a-z0-9_]+)\s*\(", open(h).read()))
    return sorted(names)
