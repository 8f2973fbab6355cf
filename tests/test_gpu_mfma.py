"""Batched BF16 hi+lo MFMA evaluation: error against the FP64 oracle per
chain, independence of a chain's result from the rest of the batch, and
the mixed-precision batched sampler's posterior against the oracle sampler.

Stated bound (mixed precision, random test points theta ~ N(0, 0.5^2)):
relative error of the log density <= 2e-6 and normwise relative error of
the gradient <= 2e-5; the measured values are printed by the tests.
"""
import sys

import numpy as np
import pytest

import oracle as O
import stan_amd as S
from stan_amd import analyze
from stan_amd import workloads as W

pytestmark = pytest.mark.gpu

LP_TOL = 2e-6
GRAD_TOL = 2e-5


def _nanmax(a, b):
    """max that keeps a NaN (the builtin max(a, nan) returns a)."""
    return a if a != a else (b if b != b else max(a, b))


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("family", [S.LOGIT, S.NORMAL])
@pytest.mark.parametrize("N,K,C", [(100, 5, 1), (4097, 37, 7), (3000, 100, 64), (2000, 256, 130),
                                   (777, 32, 65),
                                   # 256 < K <= 512: forward + link, then the adjoint
                                   # per block of 256 columns
                                   (1500, 257, 9), (2000, 300, 130), (1000, 468, 64),
                                   (600, 512, 3)])
def test_mfma_batch_against_oracle(family, N, K, C):
    d = W.logistic(N, K, 1.0 / np.sqrt(K), 0.5, seed=N + K) if family == S.LOGIT else \
        W.normal(N, K, seed=N + K)
    m = S.GlmModel(family, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)
    ref = O.GlmData(family, d["X"], d["y"], prior=d["prior"])
    th = W.random_thetas(m.D, C, seed=K + C) * 0.5
    lp, g = m.log_prob_grad_batch(th)
    worst_lp = worst_g = 0.0
    for c in range(C):
        st, lp_r, g_r = ref.lpgrad(th[c], mode=1)
        assert st == 0
        worst_lp = _nanmax(worst_lp, abs(lp[c] - lp_r) / abs(lp_r))
        worst_g = _nanmax(worst_g, np.linalg.norm(g[c] - g_r) / np.linalg.norm(g_r))
    _say(f"mfma family={family} N={N} K={K} C={C}: max rel lp err {worst_lp:.2e}, "
         f"max normwise grad err {worst_g:.2e}")
    assert worst_lp < LP_TOL and worst_g < GRAD_TOL
    # the single-chain FP64 path of the same context is unchanged
    lp64, g64 = m.log_prob_grad(th[0])
    st, lp_r, g_r = ref.lpgrad(th[0], mode=1)
    assert abs(lp64 - lp_r) <= 1e-12 * abs(lp_r)
    m.close()


def test_mfma_chain_independent_of_batch():
    d = W.logistic(5000, 64, 0.125, 0.5, seed=3)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)
    th = W.random_thetas(m.D, 130, seed=4) * 0.5
    lp_all, g_all = m.log_prob_grad_batch(th)
    for c in (0, 17, 64, 129):
        lp1, g1 = m.log_prob_grad_batch(th[c:c + 1])
        assert lp1[0] == lp_all[c]
        np.testing.assert_array_equal(g1[0], g_all[c])
    m.close()


def test_mfma_wide_chain_independent_of_batch():
    """K = 468: the two-launch path gives a chain the same result alone as
    in a batch of 130 (64- and 256-chain padding, two and four waves)."""
    d = W.logistic(3000, 468, 1.0 / np.sqrt(468), 0.5, seed=8)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)
    th = W.random_thetas(m.D, 130, seed=9) * 0.5
    lp_all, g_all = m.log_prob_grad_batch(th)
    for c in (0, 63, 64, 129):
        lp1, g1 = m.log_prob_grad_batch(th[c:c + 1])
        assert lp1[0] == lp_all[c]
        np.testing.assert_array_equal(g1[0], g_all[c])
    m.close()


def test_mfma_rejects_k_above_512():
    d = W.logistic(64, 513, 0.05, 0.5, seed=1)
    with pytest.raises(Exception, match="K <= 512"):
        S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)


def test_mfma_domain_error_is_per_chain():
    d = W.normal(500, 8, seed=2)
    m = S.GlmModel(S.NORMAL, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)
    th = W.random_thetas(m.D, 3, seed=1) * 0.3
    th[1, -1] = 1000.0  # sigma = inf
    lp, g = m.log_prob_grad_batch(th)
    assert np.isfinite(lp[0]) and np.isfinite(lp[2])
    assert lp[1] == -np.inf
    m.close()


def test_mixed_batch_sampler_posterior():
    d = W.logistic(4000, 8, 1.0 / np.sqrt(8), 0.5, seed=41)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)
    res = S.hmc_nuts_diag_e_adapt_batch(m, 16, grad_precision=S.BF16X2, random_seed=5, chain=1,
                                        num_warmup=400, num_samples=600)
    assert np.all(res["status"] == 0), res["status"]
    dev = np.transpose(res["draws"][:, :, 7:], (1, 0, 2))
    ref = O.GlmData(O.LOGIT, d["X"], d["y"], prior=d["prior"])
    f = O.model_fn("orc_glm_model_cb")
    ctx = ref.ctx(mode=0)
    cpu = []
    for chain in range(1, 5):
        rc, dr, _, _, _ = O.hmc_nuts_diag_e_adapt(f, ctx, m.D, seed=777, chain=chain,
                                                  num_warmup=400, num_samples=1000)
        assert rc == 0
        cpu.append(dr[:, 7:])
    cpu = np.stack(cpu, 1)
    for p in range(m.D):
        se = np.hypot(analyze.mcse_mean(dev[:, :, p]), analyze.mcse_mean(cpu[:, :, p]))
        assert abs(dev[:, :, p].mean() - cpu[:, :, p].mean()) < 4 * se, p
        se_sd = np.hypot(analyze.mcse_sd(dev[:, :, p]), analyze.mcse_sd(cpu[:, :, p]))
        assert abs(dev[:, :, p].std() - cpu[:, :, p].std()) < 4 * se_sd, p
    _say(f"mixed batch: min bulk ESS {analyze.min_bulk_ess(dev):.0f} over 16 chains, "
         f"steps {res['steps']}, {res['seconds']:.2f} s")
    m.close()
