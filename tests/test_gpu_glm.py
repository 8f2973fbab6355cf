"""GPU parity of the fused FP64 GLM pass against the oracle.

Bar (FP64 drop-in): log density and gradient within 1e-12 of the
compensated oracle, relative, with the gradient measured normwise and per
component against the floor eps * sum_n |x_nk r_n| (summation order is the
only legitimate source of difference).
"""
import concurrent.futures as cf
import sys

import numpy as np
import pytest

import oracle as O
import stan_amd as S
from stan_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _nanmax(a, b):
    """max that keeps a NaN (the builtin max(a, nan) returns a)."""
    return a if a != a else (b if b != b else max(a, b))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(model, ref, theta, propto=True, jacobian=True, tol=TOL):
    lp, g = model.log_prob_grad(theta, propto=propto, jacobian=jacobian)
    st, lp_ref, g_ref = ref.lpgrad(theta, mode=1, propto=propto, jacobian=jacobian)
    assert st == 0
    assert abs(lp - lp_ref) <= tol * max(1.0, abs(lp_ref)), (lp, lp_ref)
    err = np.linalg.norm(g - g_ref)
    assert err <= tol * max(1.0, np.linalg.norm(g_ref)), (err, np.linalg.norm(g_ref))
    return lp, g


@pytest.mark.parametrize("N,K", [(1, 1), (63, 3), (64, 4), (65, 5), (1000, 37),
                                 (4097, 100), (3000, 256), (2000, 512), (777, 33)])
def test_logistic_parity_shapes(N, K):
    _say(f"logistic parity N={N} K={K}")
    d = W.logistic(N, K, 1.0 / np.sqrt(K), 0.5, seed=N + K)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    ref = O.GlmData(O.LOGIT, d["X"], d["y"], prior=d["prior"])
    for th in W.random_thetas(m.D, 3, seed=K):
        _check(m, ref, th)
        _check(m, ref, th, propto=False, jacobian=False)
    m.close()


@pytest.mark.parametrize("N,K", [(1, 1), (65, 7), (5000, 64), (2000, 256), (100, 300)])
def test_normal_parity_shapes(N, K):
    _say(f"normal parity N={N} K={K}")
    d = W.normal(N, K, seed=N * 3 + K)
    for prior in [d["prior"], (5.0, 1.0, 2.0, 1.0)]:
        m = S.GlmModel(S.NORMAL, d["X"], d["y"], prior=prior)
        ref = O.GlmData(O.NORMAL, d["X"], d["y"], prior=prior)
        for th in W.random_thetas(m.D, 3, seed=K + 1) * 0.5:
            _check(m, ref, th)
            _check(m, ref, th, propto=False, jacobian=False)
        m.close()


@pytest.mark.parametrize("family", [S.LOGIT, S.NORMAL])
@pytest.mark.parametrize("lanes,cap", [(1, 32), (2, 24), (4, 28), (4, 12), (8, 16), (8, 8),
                                       (16, 4), (32, 4), (64, 4), (2, 32), (1, 20)])
@pytest.mark.parametrize("prefetch", [0, 1])
def test_every_kernel_variant(family, lanes, cap, prefetch):
    """Each layout / register capacity / prefetch variant, with an odd block
    count so that waves own different numbers of tiles."""
    K = 45 if lanes * cap >= 45 else lanes * cap
    N = 3001
    _say(f"variant family={family} L={lanes} cap={cap} pf={prefetch} K={K}")
    d = W.logistic(N, K, 0.3, 0.5, seed=K) if family == S.LOGIT else W.normal(N, K, seed=K)
    m = S.GlmModel(family, d["X"], d["y"], prior=d["prior"],
                   tuning=dict(lanes_per_row=lanes, slots_cap=cap, prefetch=prefetch,
                               blocks=7))
    cfg = m.kernel_config()
    assert cfg["lanes_per_row"] == lanes and cfg["prefetch"] == prefetch
    ref = O.GlmData(family, d["X"], d["y"], prior=d["prior"])
    for th in W.random_thetas(m.D, 2, seed=lanes + cap) * 0.5:
        _check(m, ref, th)
    m.close()


def test_autotuned_configuration_parity():
    d = W.logistic(60000, 100, 0.1, 0.5, seed=9)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    log = m.tuning_log()
    _say(log)
    assert "chosen:" in log
    ref = O.GlmData(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    for th in W.random_thetas(m.D, 3, seed=1):
        _check(m, ref, th)
    m.close()


def test_empty_data_gives_prior_only():
    X = np.zeros((0, 3))
    m = S.GlmModel(S.LOGIT, X, np.zeros(0, dtype=np.int32), prior=(5.0, 2.0, 2.0, 1.0))
    th = np.array([0.5, -1.0, 2.0, 0.25])
    lp, g = m.log_prob_grad(th)
    assert lp == pytest.approx(-0.5 * (0.1 ** 2 + 0.5 ** 2 + 1.0 + 0.125 ** 2), rel=1e-14)
    np.testing.assert_allclose(g, -th / np.array([25.0, 4.0, 4.0, 4.0]), rtol=1e-14)
    m.close()


def test_bitwise_repeatable_and_log_prob_consistent():
    d = W.logistic(20000, 50, 0.2, 0.5, seed=3)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    th = W.random_thetas(m.D, 1, seed=5)[0]
    lp1, g1 = m.log_prob_grad(th)
    lp2, g2 = m.log_prob_grad(th)
    assert lp1 == lp2 and np.array_equal(g1, g2)
    assert m.log_prob(th, propto=True, jacobian=True) == lp1
    lps, gs = m.log_prob_grad_batch(np.stack([th, th * 0.5]))
    assert lps[0] == lp1 and np.array_equal(gs[0], g1)
    m.close()


def test_domain_errors():
    d = W.normal(100, 3, seed=1)
    m = S.GlmModel(S.NORMAL, d["X"], d["y"], prior=d["prior"])
    with pytest.raises(S.DomainError):
        m.log_prob_grad(np.array([0.0, 0.0, 0.0, 0.0, 1000.0]))  # sigma = inf
    with pytest.raises(S.DomainError):
        m.log_prob_grad(np.array([np.nan, 0.0, 0.0, 0.0, 0.0]))
    # the context stays usable after a domain error
    lp, _ = m.log_prob_grad(np.zeros(5))
    assert np.isfinite(lp)
    m.close()


def test_large_logistic_parity_100_points():
    """Plain drop-in at its full size: N=1e6, K=100, X ~ N(0, 1/K), 100 test
    points theta ~ N(0, 1), FP64, within 1e-12 of the oracle."""
    _say("generating N=1e6 K=100 logistic data")
    d = W.logistic_large()
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    ref = O.GlmData(O.LOGIT, d["X"], d["y"], prior=d["prior"])
    thetas = W.random_thetas(m.D, 100, seed=2024)
    _say(f"kernel config {m.kernel_config()}")
    gpu = [m.log_prob_grad(th) for th in thetas]
    _say("GPU evaluations done; running oracle on host threads")
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        refs = list(ex.map(lambda th: ref.lpgrad(th, mode=1), thetas))
    worst_lp = worst_g = worst_c = 0.0
    Xabs = np.abs(d["X"])
    for (lp, g), (st, lp_ref, g_ref), th in zip(gpu, refs, thetas):
        assert st == 0
        worst_lp = _nanmax(worst_lp, abs(lp - lp_ref) / abs(lp_ref))
        worst_g = _nanmax(worst_g, np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref))
    # per-component floor eps * sum |x_nk r_n| at the first point
    th = thetas[0]
    eta = d["X"] @ th[1:] + th[0]
    r = d["y"] - 1.0 / (1.0 + np.exp(-eta))
    floor = np.concatenate([[np.abs(r).sum()], Xabs.T @ np.abs(r)]) * 2.0 ** -52
    worst_c = np.max(np.abs(gpu[0][1] - refs[0][2]) / (floor + np.abs(refs[0][2]) * 1e-16))
    _say(f"N=1e6 K=100: max rel lp err {worst_lp:.3e}, max normwise grad err {worst_g:.3e}, "
         f"max component err / (eps sum|x r|) {worst_c:.2f}")
    assert worst_lp < TOL
    assert worst_g < TOL
    m.close()


def test_autotuned_contexts_are_bitwise_equal():
    """The autotuner's layout and block count set the FP64 summation order;
    its first decision per device and shape is reused, so two autotuned
    contexts of one data set give bitwise-equal results."""
    print("autotune determinism", flush=True)
    d = W.logistic(300_000, 40, 1.0 / np.sqrt(40), 0.5, seed=404)
    tn = dict(autotune=1)
    a = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"], tuning=tn)
    b = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"], tuning=tn)
    assert b.tuning_log().startswith("cached:"), b.tuning_log()
    assert a.kernel_config() == b.kernel_config()
    for th in W.random_thetas(a.D, 3, seed=12):
        la, ga = a.log_prob_grad(th)
        lb, gb = b.log_prob_grad(th)
        assert la == lb
        np.testing.assert_array_equal(ga, gb)
    a.close()
    b.close()
