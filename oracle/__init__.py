"""CPU oracle for the GLM gradient + NUTS hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package, and only as the checker.  The product
(``stan_amd``) never imports, links or calls it.

The C sources restate the reference algorithms file by file; citations are
in ``oracle.h``.  This module is a thin ctypes wrapper around
``oracle/liboracle.so`` (built by ``make -C oracle``).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

LOGIT, NORMAL, HIER = 0, 1, 2


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 2 + [C.POINTER(C.c_uint32)]
        L.orc_glm_lpgrad.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                     C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double),
                                     C.c_void_p]
        L.orc_glm_lpgrad.restype = C.c_int
        L.orc_glm_num_params.argtypes = [C.c_int, C.c_int64, C.c_int64]
        L.orc_glm_num_params.restype = C.c_int64
        L.orc_hmc_nuts_diag_e_adapt.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                C.c_void_p, C.POINTER(C.c_int64),
                                                C.POINTER(C.c_double), C.c_void_p,
                                                C.POINTER(C.c_int64)]
        L.orc_hmc_nuts_diag_e_adapt.restype = C.c_int
        L.orc_hmc_nuts_dense_e_adapt.argtypes = L.orc_hmc_nuts_diag_e_adapt.argtypes
        L.orc_hmc_nuts_dense_e_adapt.restype = C.c_int
        for nm in ("orc_hmc_nuts_diag_e", "orc_hmc_nuts_dense_e"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                       C.c_void_p, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64)]
            getattr(L, nm).restype = C.c_int
        L.orc_cholesky.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_cholesky.restype = C.c_int
        L.orc_leapfrog.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
        L.orc_leapfrog.restype = C.c_int
        L.orc_hamiltonian.argtypes = [C.c_void_p]
        L.orc_hamiltonian.restype = C.c_double
        L.orc_build_tree.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_double, C.c_double,
                                     C.POINTER(C.c_int), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_build_tree.restype = C.c_int
        L.orc_learn_stepsize.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_double]
        L.orc_rng_init.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
        L.orc_rng_uniform.argtypes = [C.c_void_p]
        L.orc_rng_uniform.restype = C.c_double
        L.orc_rng_normal.argtypes = [C.c_void_p]
        L.orc_rng_normal.restype = C.c_double
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Prior(C.Structure):
    _fields_ = [("alpha_sd", C.c_double), ("beta_sd", C.c_double),
                ("sigma_sd", C.c_double), ("tau_sd", C.c_double),
                ("sigma_mu", C.c_double)]


class GlmCtx(C.Structure):
    _fields_ = [("family", C.c_int), ("mode", C.c_int), ("X", C.c_void_p),
                ("y", C.c_void_p), ("N", C.c_int64), ("K", C.c_int64),
                ("J", C.c_int64), ("offsets", C.c_void_p), ("prior", Prior)]


class Point(C.Structure):
    _fields_ = [("D", C.c_int), ("q", C.c_void_p), ("p", C.c_void_p),
                ("g", C.c_void_p), ("minv", C.c_void_p), ("V", C.c_double),
                ("dense", C.c_int)]


class Rng(C.Structure):
    _fields_ = [("key", C.c_uint32 * 2), ("counter", C.c_uint64)]


class NutsArgs(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("chain", C.c_uint32), ("init_radius", C.c_double),
                ("num_warmup", C.c_int), ("num_samples", C.c_int), ("num_thin", C.c_int),
                ("save_warmup", C.c_int), ("stepsize", C.c_double),
                ("stepsize_jitter", C.c_double), ("max_depth", C.c_int),
                ("delta", C.c_double), ("gamma", C.c_double), ("kappa", C.c_double),
                ("t0", C.c_double), ("init_buffer", C.c_uint), ("term_buffer", C.c_uint),
                ("window", C.c_uint), ("init", C.c_void_p), ("init_inv_metric", C.c_void_p)]


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def rng_stream(seed, chain, kinds):
    """Draw a sequence of variates; kinds is a string of 'u' / 'n'."""
    r = Rng()
    lib().orc_rng_init(C.byref(r), seed, chain)
    out = []
    for k in kinds:
        out.append(lib().orc_rng_uniform(C.byref(r)) if k == "u"
                   else lib().orc_rng_normal(C.byref(r)))
    return np.array(out)


def num_params(family, K, J=0):
    return lib().orc_glm_num_params(family, K, J)


class GlmData:
    """Owns contiguous copies of one GLM data set for the oracle."""

    def __init__(self, family, X, y, offsets=None, prior=(5.0, 2.0, 2.0, 1.0)):
        self.family = family
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.N, self.K = self.X.shape
        if family == NORMAL:
            self.y = np.ascontiguousarray(y, dtype=np.float64)
        else:
            self.y = np.ascontiguousarray(y, dtype=np.int32)
        self.offsets = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.int64)
        self.J = 0 if offsets is None else len(offsets) - 1
        self.prior = Prior(*prior)
        self.D = num_params(family, self.K, self.J)

    def lpgrad(self, theta, mode=1, propto=True, jacobian=True):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        g = np.zeros(self.D)
        lp = C.c_double(0.0)
        st = lib().orc_glm_lpgrad(self.family, mode, _ptr(self.X), _ptr(self.y), self.N,
                                  self.K, _ptr(self.offsets), self.J, C.byref(self.prior),
                                  _ptr(theta), int(propto), int(jacobian), C.byref(lp),
                                  _ptr(g))
        return st, lp.value, g

    def ctx(self, mode=1):
        c = GlmCtx(self.family, mode, _ptr(self.X).value, _ptr(self.y).value, self.N,
                   self.K, self.J,
                   None if self.offsets is None else _ptr(self.offsets).value, self.prior)
        c._keep = self
        return c


def model_fn(name):
    """Function pointer (as c_void_p) of a built-in C model callback."""
    f = getattr(lib(), name)
    return C.cast(f, C.c_void_p)


def default_args(**kw):
    a = dict(seed=1234, chain=1, init_radius=2.0, num_warmup=1000, num_samples=1000,
             num_thin=1, save_warmup=0, stepsize=1.0, stepsize_jitter=0.0, max_depth=10,
             delta=0.8, gamma=0.05, kappa=0.75, t0=10.0, init_buffer=75, term_buffer=50,
             window=25)
    a.update(kw)
    return a


def _run(entry, fn_ptr, ctx, D, init, init_inv_metric, metric_size, adapt, kw):
    a = default_args(**kw)
    init_a = None if init is None else np.ascontiguousarray(init, dtype=np.float64)
    minv_a = None if init_inv_metric is None else np.ascontiguousarray(init_inv_metric,
                                                                        dtype=np.float64)
    args = NutsArgs(a["seed"], a["chain"], a["init_radius"], a["num_warmup"],
                    a["num_samples"], a["num_thin"], a["save_warmup"], a["stepsize"],
                    a["stepsize_jitter"], a["max_depth"], a["delta"], a["gamma"],
                    a["kappa"], a["t0"], a["init_buffer"], a["term_buffer"], a["window"],
                    None if init_a is None else _ptr(init_a).value,
                    None if minv_a is None else _ptr(minv_a).value)
    max_rows = (a["num_warmup"] if a["save_warmup"] else 0) + a["num_samples"]
    draws = np.zeros((max(max_rows, 1), 7 + D))
    rows = C.c_int64(0)
    eps = C.c_double(0.0)
    minv = np.zeros(metric_size)
    ng = C.c_int64(0)
    cx = C.byref(ctx) if ctx is not None else None
    if adapt:
        rc = getattr(lib(), entry)(fn_ptr, cx, D, C.byref(args), _ptr(draws), C.byref(rows),
                                   C.byref(eps), _ptr(minv), C.byref(ng))
        return rc, draws[: rows.value], eps.value, minv, ng.value
    rc = getattr(lib(), entry)(fn_ptr, cx, D, C.byref(args), _ptr(draws), C.byref(rows),
                               C.byref(ng))
    return rc, draws[: rows.value], ng.value


def hmc_nuts_diag_e_adapt(fn_ptr, ctx, D, init=None, init_inv_metric=None, **kw):
    """Run the oracle sampler.  Returns (rc, draws[rows, 7+D], stepsize,
    inv_metric, n_grad)."""
    return _run("orc_hmc_nuts_diag_e_adapt", fn_ptr, ctx, D, init, init_inv_metric, D, True,
                kw)


def hmc_nuts_dense_e_adapt(fn_ptr, ctx, D, init=None, init_inv_metric=None, **kw):
    """Dense-metric adaptive sampler.  Returns (rc, draws, stepsize,
    inv_metric[D, D], n_grad)."""
    rc, d, eps, m, ng = _run("orc_hmc_nuts_dense_e_adapt", fn_ptr, ctx, D, init,
                             init_inv_metric, D * D, True, kw)
    return rc, d, eps, m.reshape(D, D), ng


def hmc_nuts_diag_e(fn_ptr, ctx, D, init=None, init_inv_metric=None, **kw):
    """Non-adaptive diag_e service.  Returns (rc, draws, n_grad)."""
    return _run("orc_hmc_nuts_diag_e", fn_ptr, ctx, D, init, init_inv_metric, D, False, kw)


def hmc_nuts_dense_e(fn_ptr, ctx, D, init=None, init_inv_metric=None, **kw):
    """Non-adaptive dense_e service.  Returns (rc, draws, n_grad)."""
    return _run("orc_hmc_nuts_dense_e", fn_ptr, ctx, D, init, init_inv_metric, D * D, False,
                kw)


def cholesky(A):
    """Lower Cholesky factor of a symmetric positive-definite matrix, or
    None if it is not positive definite."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    L = np.zeros_like(A)
    return L if lib().orc_cholesky(_ptr(A), A.shape[0], _ptr(L)) == 0 else None
