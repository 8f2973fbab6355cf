"""stan_amd: MI355X-native GLM log-density/gradient engine + NUTS sampler.

Python side of the C-ABI in ``include/stanx.h``: plain ctypes plumbing over
``stan_amd/libstanx.so`` (HIP kernels for gfx950 + the C++ sampler
services).  Importing this package never falls back to CPU code: if the
library is missing, ``lib()`` raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstanx.so")

SX_OK, SX_EDOMAIN, SX_EINVAL, SX_EDEVICE, SX_ENOMEM = 0, 1, 2, 3, 4
LOGIT, NORMAL, HIER_LOGIT = 0, 1, 2
FP64, FP32, BF16X2 = 0, 1, 2

# every entry point declared in include/stanx.h
EXPORTED = [
    "sx_glm_create", "sx_glm_create_ex", "sx_glm_tuning_log", "sx_glm_destroy", "sx_glm_num_params", "sx_glm_family",
    "sx_glm_dims",
    "sx_last_error", "sx_glm_logp_grad", "sx_glm_logp", "sx_glm_logp_grad_batch",
    "sx_glm_time_logp_grad", "sx_glm_kernel_config", "sx_hmc_nuts_diag_e_adapt",
    "sx_hmc_nuts_diag_e_adapt_cb", "sx_hmc_nuts_diag_e", "sx_hmc_nuts_diag_e_cb",
    "sx_nuts_batch_create", "sx_server_create", "sx_server_footprint", "sx_server_submit",
    "sx_server_info", "sx_server_wait", "sx_server_results", "sx_server_release",
    "sx_server_destroy", "sx_analyze_ess", "sx_hmc_nuts_diag_e_adapt_csv",
    "sx_hmc_nuts_diag_e_adapt_cb_csv", "sx_glm_create_json", "sx_json_var", "sx_nuts_batch_run",
    "sx_nuts_batch_rows_cap", "sx_nuts_batch_results", "sx_nuts_batch_destroy",
    "sx_hmc_nuts_dense_e_adapt", "sx_hmc_nuts_dense_e_adapt_cb", "sx_hmc_nuts_dense_e_adapt_csv",
    "sx_hmc_nuts_dense_e", "sx_hmc_nuts_dense_e_cb", "sx_hmc_nuts_dense_e_adapt_cb_csv",
    "sx_hmc_nuts_diag_e_adapt_files", "sx_hmc_nuts_diag_e_adapt_cb_files",
    "sx_hmc_nuts_diag_e_files", "sx_device_copy_bandwidth", "sx_glm_logp_grad_noalloc",
]

_lib = None


class StanxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class DomainError(StanxError, ValueError):
    """The reference's std::domain_error: the density cannot be evaluated."""


class Prior(C.Structure):
    _fields_ = [("alpha_sd", C.c_double), ("beta_sd", C.c_double),
                ("sigma_sd", C.c_double), ("tau_sd", C.c_double),
                ("sigma_mu", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("lanes_per_row", C.c_int), ("slots_cap", C.c_int), ("prefetch", C.c_int),
                ("blocks", C.c_int), ("autotune", C.c_int), ("hier_chunk", C.c_int),
                ("hier_kernel", C.c_int)]


class NutsConfig(C.Structure):
    _fields_ = [("random_seed", C.c_uint32), ("chain", C.c_uint32),
                ("init_radius", C.c_double), ("num_warmup", C.c_int32),
                ("num_samples", C.c_int32), ("num_thin", C.c_int32),
                ("save_warmup", C.c_int32), ("refresh", C.c_int32),
                ("stepsize", C.c_double), ("stepsize_jitter", C.c_double),
                ("max_depth", C.c_int32), ("delta", C.c_double), ("gamma", C.c_double),
                ("kappa", C.c_double), ("t0", C.c_double), ("init_buffer", C.c_uint32),
                ("term_buffer", C.c_uint32), ("window", C.c_uint32)]


class OutputFiles(C.Structure):
    _fields_ = [("sample_csv", C.c_char_p), ("diagnostic_csv", C.c_char_p),
                ("metric_json", C.c_char_p), ("sig_figs", C.c_int)]


class JobSpec(C.Structure):
    _fields_ = [("family", C.c_int), ("X", C.c_void_p), ("y", C.c_void_p), ("N", C.c_int64),
                ("K", C.c_int64), ("group_offsets", C.c_void_p), ("J", C.c_int64),
                ("prior", Prior), ("chains", C.c_int64), ("grad_precision", C.c_int),
                ("cfg", NutsConfig), ("target_min_ess", C.c_double),
                ("max_samples", C.c_int64)]


class JobInfo(C.Structure):
    _fields_ = [("state", C.c_int), ("rc", C.c_int), ("footprint_bytes", C.c_int64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("n_grad", C.c_int64), ("steps", C.c_int64), ("D", C.c_int64),
                ("chains", C.c_int64), ("rows_cap", C.c_int64),
                ("sampling_iterations", C.c_int64), ("extensions", C.c_int),
                ("min_ess", C.c_double), ("grad_precision", C.c_int)]


class NutsResult(C.Structure):
    _fields_ = [("draws", C.c_void_p), ("rows", C.c_int64), ("stepsize", C.c_double),
                ("inv_metric", C.c_void_p), ("n_grad", C.c_int64),
                ("warmup_seconds", C.c_double), ("sampling_seconds", C.c_double)]


LPGRAD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int,
                        C.POINTER(C.c_double), C.POINTER(C.c_double))


def lib():
    """Load libstanx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, i64, dp = C.c_void_p, C.c_int64, C.POINTER(C.c_double)
        L.sx_glm_create.argtypes = [C.POINTER(vp), C.c_int, vp, vp, i64, i64, vp, i64,
                                    C.c_int, C.POINTER(Prior), C.c_int]
        L.sx_glm_create_ex.argtypes = [C.POINTER(vp), C.c_int, vp, vp, i64, i64, vp, i64,
                                       C.c_int, C.POINTER(Prior), C.c_int, C.POINTER(Tuning)]
        L.sx_glm_tuning_log.argtypes = [vp]
        L.sx_glm_tuning_log.restype = C.c_char_p
        L.sx_glm_destroy.argtypes = [vp]
        L.sx_glm_destroy.restype = None
        L.sx_glm_num_params.argtypes = [vp, C.POINTER(i64)]
        L.sx_glm_family.argtypes = [vp]
        L.sx_glm_dims.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
        L.sx_last_error.restype = C.c_char_p
        L.sx_glm_logp_grad.argtypes = [vp, vp, C.c_int, C.c_int, dp, vp]
        L.sx_glm_logp_grad_noalloc.argtypes = [vp, vp, C.c_int, C.c_int, vp, vp]
        L.sx_glm_logp.argtypes = [vp, vp, C.c_int, C.c_int, dp]
        L.sx_glm_logp_grad_batch.argtypes = [vp, vp, i64, C.c_int, C.c_int, vp, vp]
        L.sx_glm_time_logp_grad.argtypes = [vp, vp, i64, C.c_int, C.c_int, dp, dp]
        L.sx_glm_kernel_config.argtypes = [vp] + [C.POINTER(C.c_int)] * 5
        L.sx_hmc_nuts_diag_e_adapt.argtypes = [vp, C.POINTER(NutsConfig), vp, vp,
                                               C.POINTER(NutsResult)]
        L.sx_hmc_nuts_diag_e_adapt_cb.argtypes = [vp, vp, C.c_int, C.POINTER(NutsConfig),
                                                  vp, vp, C.POINTER(NutsResult)]
        L.sx_hmc_nuts_diag_e.argtypes = [vp, C.POINTER(NutsConfig), vp, vp,
                                         C.POINTER(NutsResult)]
        L.sx_hmc_nuts_diag_e_cb.argtypes = [vp, vp, C.c_int, C.POINTER(NutsConfig),
                                            vp, vp, C.POINTER(NutsResult)]
        L.sx_hmc_nuts_diag_e_adapt_csv.argtypes = [vp, C.POINTER(NutsConfig), vp, vp,
                                                   C.c_char_p, C.c_int, C.POINTER(NutsResult)]
        L.sx_hmc_nuts_diag_e_adapt_cb_csv.argtypes = [vp, vp, C.c_int, C.POINTER(NutsConfig),
                                                      vp, vp, C.c_char_p, C.c_int,
                                                      C.POINTER(NutsResult)]
        L.sx_hmc_nuts_diag_e_adapt_files.argtypes = [vp, C.POINTER(NutsConfig), vp, vp,
                                                     C.POINTER(OutputFiles),
                                                     C.POINTER(NutsResult)]
        L.sx_hmc_nuts_diag_e_adapt_cb_files.argtypes = [vp, vp, C.c_int, C.POINTER(NutsConfig),
                                                        vp, vp, C.POINTER(OutputFiles),
                                                        C.POINTER(NutsResult)]
        L.sx_hmc_nuts_diag_e_files.argtypes = L.sx_hmc_nuts_diag_e_adapt_files.argtypes
        L.sx_device_copy_bandwidth.argtypes = [C.c_int, i64, C.c_int, dp, dp]
        L.sx_hmc_nuts_dense_e_adapt.argtypes = L.sx_hmc_nuts_diag_e_adapt.argtypes
        L.sx_hmc_nuts_dense_e_adapt_cb.argtypes = L.sx_hmc_nuts_diag_e_adapt_cb.argtypes
        L.sx_hmc_nuts_dense_e_adapt_csv.argtypes = L.sx_hmc_nuts_diag_e_adapt_csv.argtypes
        L.sx_hmc_nuts_dense_e_adapt_cb_csv.argtypes = L.sx_hmc_nuts_diag_e_adapt_cb_csv.argtypes
        L.sx_hmc_nuts_dense_e.argtypes = L.sx_hmc_nuts_diag_e.argtypes
        L.sx_hmc_nuts_dense_e_cb.argtypes = L.sx_hmc_nuts_diag_e_cb.argtypes
        L.sx_glm_create_json.argtypes = [C.POINTER(vp), C.c_char_p, C.c_int, C.c_int,
                                         C.POINTER(Prior), C.c_int]
        L.sx_json_var.argtypes = [C.c_char_p, C.c_char_p, vp, i64, C.POINTER(i64), C.c_int,
                                  C.POINTER(C.c_int), C.POINTER(i64)]
        L.sx_analyze_ess.argtypes = [vp, i64, i64, C.c_int, dp]
        L.sx_server_create.argtypes = [C.POINTER(vp), C.c_int, i64, C.c_int]
        L.sx_server_footprint.argtypes = [C.POINTER(JobSpec), C.POINTER(i64)]
        L.sx_server_submit.argtypes = [vp, C.POINTER(JobSpec), C.POINTER(i64)]
        L.sx_server_info.argtypes = [vp, i64, C.POINTER(JobInfo)]
        L.sx_server_wait.argtypes = [vp, i64, C.c_double]
        L.sx_server_results.argtypes = [vp, i64, vp, vp, vp]
        L.sx_server_release.argtypes = [vp, i64]
        L.sx_server_destroy.argtypes = [vp]
        L.sx_server_destroy.restype = None
        L.sx_nuts_batch_create.argtypes = [C.POINTER(vp), vp, C.POINTER(NutsConfig), i64,
                                           C.c_int, vp, vp]
        L.sx_nuts_batch_run.argtypes = [vp, i64, C.POINTER(i64), C.POINTER(i64)]
        L.sx_nuts_batch_rows_cap.argtypes = [vp]
        L.sx_nuts_batch_rows_cap.restype = i64
        L.sx_nuts_batch_results.argtypes = [vp, vp, vp, vp, vp, vp, vp, dp, C.POINTER(i64)]
        L.sx_nuts_batch_destroy.argtypes = [vp]
        L.sx_nuts_batch_destroy.restype = None
        _lib = L
    return _lib


def device_copy_bandwidth(nbytes=1 << 31, iters=20, device=0):
    """Measured copy bandwidth of the device (GB/s, read + write) and the
    milliseconds per copy of `nbytes`."""
    gbs, ms = C.c_double(), C.c_double()
    _check(lib().sx_device_copy_bandwidth(device, int(nbytes), int(iters), C.byref(gbs),
                                          C.byref(ms)))
    return gbs.value, ms.value


def last_error():
    return lib().sx_last_error().decode(errors="replace")


def _check(rc):
    if rc == SX_OK:
        return
    msg = last_error()
    if rc == SX_EDOMAIN:
        raise DomainError(rc, msg)
    raise StanxError(rc, msg)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class GlmModel:
    """A GLM log density with its data resident on the GPU.

    family: LOGIT, NORMAL, or HIER_LOGIT with ``offsets`` (J + 1 CSR row
    offsets of the groups).  X is N x K, y int {0,1} (logistic) or float
    (normal).  precision: FP64; BF16X2 (LOGIT/NORMAL: batched evaluations on
    MFMA); FP32 (HIER_LOGIT: batched evaluations in FP32).  Theta layouts
    are those of include/stanx.h.
    """

    def __init__(self, family, X, y, offsets=None, precision=FP64,
                 prior=(5.0, 2.0, 2.0, 1.0), device=0, tuning=None):
        X = np.ascontiguousarray(X, dtype=np.float64)
        if X.ndim != 2:
            raise ValueError("X must be N x K")
        N, K = X.shape
        if family == NORMAL:
            y = np.ascontiguousarray(y, dtype=np.float64)
        else:
            y = np.ascontiguousarray(y, dtype=np.int32)
        if y.shape != (N,):
            raise ValueError("y must have N entries")
        off = None
        J = 0
        if offsets is not None:
            off = np.ascontiguousarray(offsets, dtype=np.int64)
            J = len(off) - 1
        self._h = C.c_void_p()
        self.family, self.N, self.K, self.J = family, N, K, J
        self.prior = Prior(*prior)
        tn = None if tuning is None else Tuning(**{**dict(lanes_per_row=0, slots_cap=0,
                                                             prefetch=-1, blocks=0,
                                                             autotune=0, hier_chunk=0,
                                                             hier_kernel=0),
                                                        **tuning})
        _check(lib().sx_glm_create_ex(C.byref(self._h), family, _ptr(X), _ptr(y), N, K,
                                      _ptr(off), J, precision, C.byref(self.prior), device,
                                      None if tn is None else C.byref(tn)))
        D = C.c_int64()
        _check(lib().sx_glm_num_params(self._h, C.byref(D)))
        self.D = D.value

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().sx_glm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def log_prob_grad(self, theta, propto=True, jacobian=True):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        g = np.zeros(self.D)
        lp = C.c_double()
        _check(lib().sx_glm_logp_grad(self._h, _ptr(theta), int(propto), int(jacobian),
                                      C.byref(lp), _ptr(g)))
        return lp.value, g

    def log_prob(self, theta, propto=False, jacobian=True):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        lp = C.c_double()
        _check(lib().sx_glm_logp(self._h, _ptr(theta), int(propto), int(jacobian),
                                 C.byref(lp)))
        return lp.value

    def log_prob_grad_batch(self, theta, propto=True, jacobian=True):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        Cn = theta.shape[0]
        lp = np.zeros(Cn)
        g = np.zeros((Cn, self.D))
        rc = lib().sx_glm_logp_grad_batch(self._h, _ptr(theta), Cn, int(propto),
                                          int(jacobian), _ptr(lp), _ptr(g))
        if rc not in (SX_OK, SX_EDOMAIN):
            _check(rc)
        return lp, g

    def time_log_prob_grad(self, theta, iters=100, warmup=10, chains=1):
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        ms, kms = C.c_double(), C.c_double()
        _check(lib().sx_glm_time_logp_grad(self._h, _ptr(theta), chains, iters, warmup,
                                           C.byref(ms), C.byref(kms)))
        return ms.value, kms.value

    def kernel_config(self):
        v = [C.c_int() for _ in range(5)]
        _check(lib().sx_glm_kernel_config(self._h, *[C.byref(x) for x in v]))
        return dict(lanes_per_row=v[0].value, slots_per_lane=v[1].value,
                    blocks=v[2].value, threads=v[3].value, prefetch=v[4].value)

    def tuning_log(self):
        return lib().sx_glm_tuning_log(self._h).decode()


def nuts_config(**kw):
    """Defaults of the reference service (src/stan/services/sample/defaults.hpp;
    delta 0.8 as in CmdStan)."""
    a = dict(random_seed=1234, chain=1, init_radius=2.0, num_warmup=1000,
             num_samples=1000, num_thin=1, save_warmup=0, refresh=0, stepsize=1.0,
             stepsize_jitter=0.0, max_depth=10, delta=0.8, gamma=0.05, kappa=0.75,
             t0=10.0, init_buffer=75, term_buffer=50, window=25)
    a.update(kw)
    return NutsConfig(**a)


def _run_service(call, D, cfg, init, inv_metric, dense=False):
    rows_cap = (cfg.num_warmup if cfg.save_warmup else 0) + cfg.num_samples
    draws = np.zeros((max(rows_cap, 1), 7 + D))
    minv = np.zeros((D, D) if dense else D)
    res = NutsResult(_ptr(draws).value, 0, 0.0, _ptr(minv).value, 0, 0.0, 0.0)
    init_a = None if init is None else np.ascontiguousarray(init, dtype=np.float64)
    minv_a = None if inv_metric is None else np.ascontiguousarray(inv_metric,
                                                                   dtype=np.float64)
    rc = call(init_a, minv_a, res)
    return dict(rc=rc, draws=draws[: res.rows], stepsize=res.stepsize, inv_metric=minv,
                n_grad=res.n_grad, warmup_seconds=res.warmup_seconds,
                sampling_seconds=res.sampling_seconds,
                error=last_error() if rc else "")


def hmc_nuts_diag_e_adapt(model, init=None, inv_metric=None, **kw):
    """Adaptive diag_e NUTS on a GlmModel (one chain, FP64 gradient on the GPU).

    Returns a dict with rc (0 / 70 / 78), draws (rows x (7 + D): lp__,
    accept_stat__, stepsize__, treedepth__, n_leapfrog__, divergent__, energy__,
    constrained parameters), stepsize, inv_metric, n_grad, timings."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e_adapt(model.handle, C.byref(cfg), _ptr(i),
                                                       _ptr(m), C.byref(r)),
        model.D, cfg, init, inv_metric)


def hmc_nuts_diag_e_adapt_cb(fn_ptr, ctx, D, init=None, inv_metric=None, **kw):
    """The same service for a model given as a C callback (function pointer
    with the sx_lpgrad_fn signature)."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e_adapt_cb(fn_ptr, ctx, D, C.byref(cfg),
                                                          _ptr(i), _ptr(m), C.byref(r)),
        D, cfg, init, inv_metric)


def hmc_nuts_diag_e_adapt_csv(model, path, sig_figs=-1, init=None, inv_metric=None, **kw):
    """hmc_nuts_diag_e_adapt writing a Stan CSV file at `path` (readable by
    stan_amd.stan_csv.read_stan_csv and CmdStan tools); returns the same dict."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e_adapt_csv(model.handle, C.byref(cfg), _ptr(i),
                                                           _ptr(m), str(path).encode(),
                                                           int(sig_figs), C.byref(r)),
        model.D, cfg, init, inv_metric)


def hmc_nuts_diag_e_adapt_cb_csv(fn_ptr, ctx, D, path, sig_figs=-1, init=None,
                                 inv_metric=None, **kw):
    """The CSV-writing service for a model given as a C callback."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e_adapt_cb_csv(fn_ptr, ctx, D, C.byref(cfg),
                                                              _ptr(i), _ptr(m),
                                                              str(path).encode(),
                                                              int(sig_figs), C.byref(r)),
        D, cfg, init, inv_metric)


def _files(sample_csv, diagnostic_csv, metric_json, sig_figs):
    enc = lambda p: None if p is None else str(p).encode()  # noqa: E731
    return OutputFiles(enc(sample_csv), enc(diagnostic_csv), enc(metric_json), int(sig_figs))


def hmc_nuts_diag_e_adapt_files(model, sample_csv=None, diagnostic_csv=None, metric_json=None,
                                sig_figs=-1, init=None, inv_metric=None, **kw):
    """hmc_nuts_diag_e_adapt with the reference's three outputs as files: the
    sample CSV, the diagnostic CSV (q, p_*, g_* per saved draw) and the
    metric JSON (adapted stepsize and inverse metric).  Paths may be None."""
    cfg = nuts_config(**kw)
    files = _files(sample_csv, diagnostic_csv, metric_json, sig_figs)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e_adapt_files(model.handle, C.byref(cfg), _ptr(i),
                                                             _ptr(m), C.byref(files),
                                                             C.byref(r)),
        model.D, cfg, init, inv_metric)


def hmc_nuts_diag_e_adapt_cb_files(fn_ptr, ctx, D, sample_csv=None, diagnostic_csv=None,
                                   metric_json=None, sig_figs=-1, init=None, inv_metric=None,
                                   **kw):
    """hmc_nuts_diag_e_adapt_files for a model given as a C callback."""
    cfg = nuts_config(**kw)
    files = _files(sample_csv, diagnostic_csv, metric_json, sig_figs)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e_adapt_cb_files(fn_ptr, ctx, D, C.byref(cfg),
                                                                _ptr(i), _ptr(m),
                                                                C.byref(files), C.byref(r)),
        D, cfg, init, inv_metric)


def hmc_nuts_diag_e_files(model, sample_csv=None, diagnostic_csv=None, sig_figs=-1, init=None,
                          inv_metric=None, **kw):
    """hmc_nuts_diag_e (no adaptation) with sample and diagnostic CSV files."""
    cfg = nuts_config(**kw)
    files = _files(sample_csv, diagnostic_csv, None, sig_figs)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e_files(model.handle, C.byref(cfg), _ptr(i),
                                                       _ptr(m), C.byref(files), C.byref(r)),
        model.D, cfg, init, inv_metric)


def hmc_nuts_diag_e(model, init=None, inv_metric=None, **kw):
    """diag_e NUTS without adaptation on a GlmModel (fixed step size and
    metric); same return dict as hmc_nuts_diag_e_adapt."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e(model.handle, C.byref(cfg), _ptr(i), _ptr(m),
                                                 C.byref(r)),
        model.D, cfg, init, inv_metric)


def hmc_nuts_diag_e_cb(fn_ptr, ctx, D, init=None, inv_metric=None, **kw):
    """hmc_nuts_diag_e for a model given as a C callback."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_diag_e_cb(fn_ptr, ctx, D, C.byref(cfg), _ptr(i),
                                                    _ptr(m), C.byref(r)),
        D, cfg, init, inv_metric)


def hmc_nuts_dense_e_adapt(model, init=None, inv_metric=None, **kw):
    """Adaptive NUTS with a dense metric on a GlmModel: inv_metric (D x D,
    identity if None) is adapted by windowed covariance estimation; the
    returned inv_metric is D x D.  Otherwise as hmc_nuts_diag_e_adapt."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_dense_e_adapt(model.handle, C.byref(cfg), _ptr(i),
                                                        _ptr(m), C.byref(r)),
        model.D, cfg, init, inv_metric, dense=True)


def hmc_nuts_dense_e_adapt_cb(fn_ptr, ctx, D, init=None, inv_metric=None, **kw):
    """hmc_nuts_dense_e_adapt for a model given as a C callback."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_dense_e_adapt_cb(fn_ptr, ctx, D, C.byref(cfg),
                                                           _ptr(i), _ptr(m), C.byref(r)),
        D, cfg, init, inv_metric, dense=True)


def hmc_nuts_dense_e_adapt_csv(model, path, sig_figs=-1, init=None, inv_metric=None, **kw):
    """hmc_nuts_dense_e_adapt writing a Stan CSV file at `path`."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_dense_e_adapt_csv(model.handle, C.byref(cfg),
                                                            _ptr(i), _ptr(m),
                                                            str(path).encode(),
                                                            int(sig_figs), C.byref(r)),
        model.D, cfg, init, inv_metric, dense=True)


def hmc_nuts_dense_e_adapt_cb_csv(fn_ptr, ctx, D, path, sig_figs=-1, init=None,
                                  inv_metric=None, **kw):
    """The dense CSV-writing service for a model given as a C callback."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_dense_e_adapt_cb_csv(fn_ptr, ctx, D, C.byref(cfg),
                                                               _ptr(i), _ptr(m),
                                                               str(path).encode(),
                                                               int(sig_figs), C.byref(r)),
        D, cfg, init, inv_metric, dense=True)


def hmc_nuts_dense_e(model, init=None, inv_metric=None, **kw):
    """Dense-metric NUTS without adaptation on a GlmModel."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_dense_e(model.handle, C.byref(cfg), _ptr(i), _ptr(m),
                                                  C.byref(r)),
        model.D, cfg, init, inv_metric, dense=True)


def hmc_nuts_dense_e_cb(fn_ptr, ctx, D, init=None, inv_metric=None, **kw):
    """hmc_nuts_dense_e for a model given as a C callback."""
    cfg = nuts_config(**kw)
    return _run_service(
        lambda i, m, r: lib().sx_hmc_nuts_dense_e_cb(fn_ptr, ctx, D, C.byref(cfg), _ptr(i),
                                                     _ptr(m), C.byref(r)),
        D, cfg, init, inv_metric, dense=True)


class NutsBatch:
    """Many chains of adaptive diag_e NUTS resident on the device (one
    GlmModel).  Chain ids are chain, chain + 1, ...; each chain draws from
    the Philox stream of the single-chain service with the same id."""

    def __init__(self, model, chains, grad_precision=FP64, inits=None, inv_metric=None,
                 **kw):
        self.model = model
        self.C = int(chains)
        self.cfg = nuts_config(**kw)
        self._h = C.c_void_p()
        init_a = None if inits is None else np.ascontiguousarray(inits, dtype=np.float64)
        minv_a = None if inv_metric is None else np.ascontiguousarray(inv_metric,
                                                                       dtype=np.float64)
        _check(lib().sx_nuts_batch_create(C.byref(self._h), model.handle, C.byref(self.cfg),
                                          self.C, grad_precision, _ptr(init_a), _ptr(minv_a)))
        self.rows_cap = lib().sx_nuts_batch_rows_cap(self._h)

    def run(self, max_steps=1 << 40):
        steps, done = C.c_int64(), C.c_int64()
        _check(lib().sx_nuts_batch_run(self._h, int(max_steps), C.byref(steps), C.byref(done)))
        return steps.value, done.value

    def results(self):
        D = self.model.D
        draws = np.zeros((self.C, max(self.rows_cap, 1), 7 + D))
        rows = np.zeros(self.C, dtype=np.int64)
        status = np.zeros(self.C, dtype=np.int32)
        eps = np.zeros(self.C)
        minv = np.zeros((self.C, D))
        ng = np.zeros(self.C, dtype=np.int64)
        secs = C.c_double()
        steps = C.c_int64()
        _check(lib().sx_nuts_batch_results(self._h, _ptr(draws), _ptr(rows), _ptr(status),
                                           _ptr(eps), _ptr(minv), _ptr(ng), C.byref(secs),
                                           C.byref(steps)))
        return dict(draws=draws[:, : self.rows_cap], rows=rows, status=status, stepsize=eps,
                    inv_metric=minv, n_grad=ng, seconds=secs.value, steps=steps.value)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().sx_nuts_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hmc_nuts_diag_e_adapt_batch(model, chains, grad_precision=FP64, inits=None,
                                inv_metric=None, **kw):
    """Run `chains` device-resident chains to completion; see NutsBatch."""
    b = NutsBatch(model, chains, grad_precision, inits, inv_metric, **kw)
    try:
        b.run()
        return b.results()
    finally:
        b.close()


SX_JOB_QUEUED, SX_JOB_RUNNING, SX_JOB_DONE, SX_JOB_FAILED = 0, 1, 2, 3
SX_ETIMEOUT = 5


def job_spec(family, X, y, chains, offsets=None, prior=(5.0, 2.0, 2.0, 1.0),
             grad_precision=BF16X2, target_min_ess=0.0, max_samples=0, **kw):
    """A serving job: data, chains, sampler configuration (nuts_config
    keywords) and the min-ESS target.  Returns (spec, arrays kept alive)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    N, K = X.shape
    y = np.ascontiguousarray(y, dtype=np.float64 if family == NORMAL else np.int32)
    off = None if offsets is None else np.ascontiguousarray(offsets, dtype=np.int64)
    J = 0 if off is None else len(off) - 1
    spec = JobSpec(family, _ptr(X).value if X.size else _ptr(np.zeros(1)).value,
                   _ptr(y).value, N, K, None if off is None else _ptr(off).value, J,
                   Prior(*prior), chains, grad_precision, nuts_config(**kw),
                   float(target_min_ess), int(max_samples))
    return spec, (X, y, off)


def job_footprint(spec):
    b = C.c_int64()
    _check(lib().sx_server_footprint(C.byref(spec), C.byref(b)))
    return b.value


class Server:
    """Request loop over posterior jobs on one GPU (admission by HBM
    footprint, device-resident chains, overlapping work, min-ESS target)."""

    def __init__(self, hbm_budget_bytes, max_active=16, device=0):
        self._h = C.c_void_p()
        _check(lib().sx_server_create(C.byref(self._h), device, int(hbm_budget_bytes),
                                      int(max_active)))
        self._jobs = {}

    def submit(self, family, X, y, chains, **kw):
        spec, keep = job_spec(family, X, y, chains, **kw)
        jid = C.c_int64()
        _check(lib().sx_server_submit(self._h, C.byref(spec), C.byref(jid)))
        self._jobs[jid.value] = spec.K + (2 if family == NORMAL else 1)
        return jid.value

    def info(self, jid):
        inf = JobInfo()
        _check(lib().sx_server_info(self._h, jid, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in JobInfo._fields_}

    def wait(self, jid, timeout=-1.0):
        rc = lib().sx_server_wait(self._h, jid, float(timeout))
        if rc == SX_ETIMEOUT:
            return False
        _check(rc)
        return True

    def results(self, jid):
        inf = self.info(jid)
        Cn, cap, D = inf["chains"], inf["rows_cap"], inf["D"]
        draws = np.zeros((Cn, max(cap, 1), 7 + D))
        rows = np.zeros(Cn, dtype=np.int64)
        status = np.zeros(Cn, dtype=np.int32)
        _check(lib().sx_server_results(self._h, jid, _ptr(draws), _ptr(rows), _ptr(status)))
        return dict(draws=draws[:, :cap], rows=rows, status=status, info=inf)

    def release(self, jid):
        _check(lib().sx_server_release(self._h, jid))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().sx_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def native_ess(chains, kind="bulk"):
    """ESS of (iterations x chains) draws by the library's C++ routines
    (kind "ess" or "bulk"), as the serving loop computes it."""
    x = np.ascontiguousarray(np.asarray(chains, dtype=np.float64).T)
    out = C.c_double()
    _check(lib().sx_analyze_ess(_ptr(x), x.shape[1], x.shape[0], 0 if kind == "ess" else 1,
                                C.byref(out)))
    return out.value


def json_var(path, name):
    """A variable of a Stan JSON data file as read by the library (numpy
    array of the file's shape)."""
    size, nd = C.c_int64(), C.c_int()
    dims = (C.c_int64 * 8)()
    _check(lib().sx_json_var(str(path).encode(), name.encode(), None, 0, dims, 8, C.byref(nd),
                             C.byref(size)))
    vals = np.zeros(size.value)
    _check(lib().sx_json_var(str(path).encode(), name.encode(), _ptr(vals), size.value, dims, 8,
                             C.byref(nd), C.byref(size)))
    shape = tuple(dims[i] for i in range(nd.value))
    return vals.reshape(shape, order="F") if shape else vals[0]


def glm_from_json(path, family, precision=FP64, prior=(5.0, 2.0, 2.0, 1.0), device=0):
    """A GlmModel whose data come from a Stan JSON data file (X, y or Y;
    a group index for HIER_LOGIT)."""
    m = GlmModel.__new__(GlmModel)
    m._h = C.c_void_p()
    m.prior = Prior(*prior)
    _check(lib().sx_glm_create_json(C.byref(m._h), str(path).encode(), family, precision,
                                    C.byref(m.prior), device))
    D = C.c_int64()
    _check(lib().sx_glm_num_params(m._h, C.byref(D)))
    N, K, J = C.c_int64(), C.c_int64(), C.c_int64()
    _check(lib().sx_glm_dims(m._h, C.byref(N), C.byref(K), C.byref(J)))
    m.family, m.N, m.K, m.J, m.D = family, N.value, K.value, J.value, D.value
    return m
