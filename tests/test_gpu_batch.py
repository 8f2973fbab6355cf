"""Device-resident batched NUTS: each chain must follow the single-chain
host service with the same chain id (same Philox stream, same phases), and
the pooled posterior must match the oracle sampler's within 4 MCSE."""
import sys

import numpy as np
import pytest

import oracle as O
import stan_amd as S
from stan_amd import analyze
from stan_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("family", [S.LOGIT, S.NORMAL])
def test_batch_chains_follow_host_service(family):
    _say(f"batch vs host family={family}")
    d = W.logistic(3000, 6, 0.4, 0.5, seed=5) if family == S.LOGIT else W.normal(3000, 6, seed=6)
    m = S.GlmModel(family, d["X"], d["y"], prior=d["prior"])
    cfg = dict(random_seed=4242, chain=1, num_warmup=150, num_samples=100, save_warmup=1,
               max_depth=8)
    res = S.hmc_nuts_diag_e_adapt_batch(m, 4, **cfg)
    assert np.all(res["status"] == 0), res["status"]
    assert np.all(res["rows"] == 250)
    fractions = []
    for c in range(4):
        host = S.hmc_nuts_diag_e_adapt(m, **{**cfg, "chain": 1 + c})
        assert host["rc"] == 0
        dev, hd = res["draws"][c], host["draws"]
        assert dev.shape == hd.shape
        # Same stream, same phases.  Device dot products are wave reductions
        # and the batched gradient uses another launch shape, so last-bit
        # differences grow slowly (fastest through accept_stat__ when the
        # early warm-up step size is unstable, then through dual averaging).
        # A logic difference (criterion, RNG order, adaptation) would move
        # tree sizes and positions by O(1) at once.
        np.testing.assert_allclose(dev[:20], hd[:20], rtol=1e-5, atol=1e-8)
        discrete = np.all(dev[:, 3:6] == hd[:, 3:6], axis=1)
        near = np.all(np.isclose(dev, hd, rtol=1e-3, atol=1e-6), axis=1)
        _say(f"chain {c}: discrete columns equal in {discrete.mean():.3f} of rows, "
             f"all columns within 1e-3 in {near.mean():.3f}")
        assert discrete[:50].all()
        fractions.append(discrete.mean())
    assert np.mean(fractions) > 0.75
    m.close()


@pytest.mark.parametrize("num_warmup", [100, 19, 300])
def test_batch_metric_windows_follow_host_plan(num_warmup):
    """The device chains close their metric windows where the host sampler
    does.  num_warmup = 100 takes the reference's reduced 15/75/10 split, in
    which no window ever closes (the metric stays at its initial value);
    19 estimates nothing; 300 closes windows at 99, 149 and 249."""
    _say(f"metric windows num_warmup={num_warmup}")
    d = W.logistic(2000, 5, 0.4, 0.5, seed=15)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    cfg = dict(random_seed=515, chain=1, num_warmup=num_warmup, num_samples=20, max_depth=8)
    res = S.hmc_nuts_diag_e_adapt_batch(m, 2, **cfg)
    assert np.all(res["status"] == 0), res["status"]
    for c in range(2):
        host = S.hmc_nuts_diag_e_adapt(m, **{**cfg, "chain": 1 + c})
        assert host["rc"] == 0
        if num_warmup < 150:
            assert np.all(host["inv_metric"] == 1.0)
            assert np.all(res["inv_metric"][c] == 1.0)
        else:
            # both estimated the metric; last-bit differences of the device
            # reductions can move a chain's path, so the estimates of the
            # same posterior variances agree only statistically
            assert not np.any(host["inv_metric"] == 1.0)
            assert not np.any(res["inv_metric"][c] == 1.0)
            np.testing.assert_allclose(res["inv_metric"][c], host["inv_metric"], rtol=0.5)
    m.close()


def test_batch_posterior_matches_oracle_sampler():
    d = W.logistic(2000, 5, 1.0 / np.sqrt(5), 0.5, seed=77)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    res = S.hmc_nuts_diag_e_adapt_batch(m, 8, random_seed=99, chain=1, num_warmup=500,
                                        num_samples=1000)
    assert np.all(res["status"] == 0)
    dev = np.transpose(res["draws"][:, :, 7:], (1, 0, 2))  # iterations x chains x params
    ref = O.GlmData(O.LOGIT, d["X"], d["y"], prior=d["prior"])
    f = O.model_fn("orc_glm_model_cb")
    ctx = ref.ctx(mode=0)
    cpu = []
    for chain in range(1, 5):
        rc, dr, _, _, _ = O.hmc_nuts_diag_e_adapt(f, ctx, m.D, seed=31337, chain=chain,
                                                  num_warmup=500, num_samples=1000)
        assert rc == 0
        cpu.append(dr[:, 7:])
    cpu = np.stack(cpu, 1)
    for p in range(m.D):
        se = np.hypot(analyze.mcse_mean(dev[:, :, p]), analyze.mcse_mean(cpu[:, :, p]))
        assert abs(dev[:, :, p].mean() - cpu[:, :, p].mean()) < 4 * se, p
        se_sd = np.hypot(analyze.mcse_sd(dev[:, :, p]), analyze.mcse_sd(cpu[:, :, p]))
        assert abs(dev[:, :, p].std() - cpu[:, :, p].std()) < 4 * se_sd, p
    _say(f"batch min bulk ESS {analyze.min_bulk_ess(dev):.0f}, steps {res['steps']}")
    m.close()


def test_batch_mix_independence():
    """A chain's draws do not depend on how many chains run beside it."""
    d = W.logistic(1000, 4, 0.5, 0.5, seed=8)
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    cfg = dict(random_seed=7, chain=3, num_warmup=100, num_samples=50)
    one = S.hmc_nuts_diag_e_adapt_batch(m, 1, **cfg)
    many = S.hmc_nuts_diag_e_adapt_batch(m, 5, **{**cfg, "chain": 1})
    # chain id 3 is index 2 of the 5-chain batch; a chain's gradient uses the
    # same launch shape and reduction order in both runs: bitwise equal
    np.testing.assert_array_equal(one["draws"][0], many["draws"][2])
    m.close()
