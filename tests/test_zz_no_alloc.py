"""The allocation-free entry point (sx_glm_logp_grad_noalloc): device-
resident theta and output, partial sums in the library's per-GPU scratch
array.  Its result must equal sx_glm_logp_grad's (same kernels, same launch
shape, so bitwise) for each family, on consecutive calls whose second uses
a smaller data set (the scratch array keeps the first call's contents), and
on a caller's stream.  A context whose partial sums exceed the array is
rejected.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import stan_amd as S
from stan_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _hip():
    """The HIP runtime the library itself links (device buffers and streams
    for the test, through the runtime's C API)."""
    S.lib()
    h = C.CDLL("libamdhip64.so.7")
    vp = C.c_void_p
    h.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    h.hipStreamCreate.argtypes = [C.POINTER(vp)]
    h.hipStreamDestroy.argtypes = [vp]
    return h


H2D, D2H = 1, 2


def _noalloc(m, theta, stream=None):
    h = _hip()
    th = np.ascontiguousarray(theta, dtype=np.float64)
    out = np.full(m.D + 2, np.nan)
    d_th, d_out = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(d_th), th.nbytes) == 0
    assert h.hipMalloc(C.byref(d_out), out.nbytes) == 0
    try:
        assert h.hipMemcpy(d_th, th.ctypes.data, th.nbytes, H2D) == 0
        assert h.hipMemcpy(d_out, out.ctypes.data, out.nbytes, H2D) == 0
        rc = S.lib().sx_glm_logp_grad_noalloc(m.handle, d_th, 1, 1, d_out, stream)
        assert rc == S.SX_OK, S.lib().sx_last_error().decode()
        assert h.hipDeviceSynchronize() == 0
        assert h.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
    finally:
        h.hipFree(d_th)
        h.hipFree(d_out)
    return out[0], out[1:m.D + 1], out[m.D + 1]


def _models():
    big = W.logistic(20_000, 60, 1.0 / np.sqrt(60), 0.5, seed=11)
    small = W.normal(3_000, 9, seed=12)
    hier = W.hier_logistic(J=40, K=6, seed=13)
    return [
        ("logistic N=20000 K=60", S.GlmModel(S.LOGIT, big["X"], big["y"], prior=big["prior"])),
        ("normal N=3000 K=9", S.GlmModel(S.NORMAL, small["X"], small["y"],
                                         prior=small["prior"])),
        ("hierarchical J=40 K=6", S.GlmModel(S.HIER_LOGIT, hier["X"], hier["y"],
                                             offsets=hier["offsets"], prior=hier["prior"])),
    ]


def test_noalloc_equals_logp_grad_on_consecutive_calls():
    models = _models()
    try:
        for name, m in models:
            for seed in (1, 2):
                th = W.random_thetas(m.D, 1, seed=seed)[0] * 0.5
                lp_ref, g_ref = m.log_prob_grad(th)
                lp, g, status = _noalloc(m, th)
                _say(f"{name} seed {seed}: lp {lp:.17g} vs {lp_ref:.17g}")
                assert status == 0.0
                assert lp == lp_ref
                np.testing.assert_array_equal(g, g_ref)
    finally:
        for _, m in models:
            m.close()


def test_noalloc_on_a_caller_stream_and_domain_status():
    d = W.logistic(5_000, 20, 0.2, 0.5, seed=21)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    h = _hip()
    stream = C.c_void_p()
    assert h.hipStreamCreate(C.byref(stream)) == 0
    try:
        th = W.random_thetas(m.D, 1, seed=5)[0]
        lp_ref, g_ref = m.log_prob_grad(th)
        lp, g, status = _noalloc(m, th, stream=stream)
        assert status == 0.0 and lp == lp_ref
        np.testing.assert_array_equal(g, g_ref)
        bad = th.copy()
        bad[1] = 1e308  # eta overflows to inf: the density cannot be evaluated
        with pytest.raises(S.DomainError):
            m.log_prob_grad(bad)
        _, _, status = _noalloc(m, bad)
        assert status != 0.0
    finally:
        h.hipStreamDestroy(stream)
        m.close()


def test_noalloc_rejects_partials_beyond_the_scratch_array():
    # blocks x (K + 3) doubles: 2048 blocks at K = 1000 need 16.4 MB > 8 MiB
    d = W.logistic(40_000, 1000, 1.0 / np.sqrt(1000), 0.5, seed=31)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"],
                   tuning=dict(lanes_per_row=64, blocks=2048, autotune=0))
    h = _hip()
    d_th, d_out = C.c_void_p(), C.c_void_p()
    assert h.hipMalloc(C.byref(d_th), 8 * m.D) == 0
    assert h.hipMalloc(C.byref(d_out), 8 * (m.D + 2)) == 0
    try:
        rc = S.lib().sx_glm_logp_grad_noalloc(m.handle, d_th, 1, 1, d_out, None)
        assert rc == S.SX_EINVAL
        assert "scratch" in S.lib().sx_last_error().decode()
    finally:
        h.hipFree(d_th)
        h.hipFree(d_out)
        m.close()
