"""The benchmark configurations at their full sizes on the GPU.

  C1  logistic N=1e4, K=50, X ~ N(0,1): FP64 parity where the link
      saturates (|eta| > 20 occurs), and 4 device chains against the oracle
      sampler (posterior means and sds within 4 MCSE).
  C2  logistic N=1e6, K=100: the BF16 hi+lo batched gradient of 8 of 256
      chains against the FP64 oracle at theta ~ N(0,1) and N(0, 0.5^2).
  C3  normal-id N=2e5, K=256, 256 chains: the same error check; the
      256-chain BF16 sampler to completion (status 0, split R-hat, pooled
      means within 4 MCSE of an FP64 device run on the same data).
  C4  hierarchical logistic J=1000, K=20, 64 chains FP32: the batched FP32
      gradient against the FP64 oracle; the 64-chain sampler (1000 warm-up
      iterations, max_depth 10) to completion (status 0, every chain in the
      typical set: lp, step size, divergences; split R-hat of the global
      parameters).
  C5  the serving stream's largest shapes (K=468, N=8e5) alone and in a
      mix: bitwise the same draws.

Mixed-precision bounds (stated per configuration; measured values are
printed): C2 and C3 gradient normwise relative error <= 1e-6 at
theta ~ N(0, 0.5^2) (the typical set of these posteriors lies well inside
that scale) and <= 3e-6 at theta ~ N(0, 1); log density <= 1e-7 relative.
C4 (FP32 X and products, FP64 across chunks): log density <= 1e-6,
gradient <= 1e-5.
"""
import sys

import numpy as np
import pytest

import oracle as O
import stan_amd as S
from stan_amd import analyze
from stan_amd import workloads as W

pytestmark = pytest.mark.gpu

GB = 1 << 30


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _rel_errors(model, ref, thetas):
    lp, g = model.log_prob_grad_batch(thetas)
    e_lp = e_g = 0.0
    for c in range(len(thetas)):
        st, lp_r, g_r = ref.lpgrad(thetas[c], mode=1)
        assert st == 0
        e_lp = max(e_lp, abs(lp[c] - lp_r) / abs(lp_r))
        e_g = max(e_g, float(np.linalg.norm(g[c] - g_r) / np.linalg.norm(g_r)))
    return e_lp, e_g


def _split_rhat(x):
    """x: iterations x chains; split R-hat (Gelman et al. 2013)."""
    n = x.shape[0] // 2
    s = np.concatenate([x[:n], x[n:2 * n]], axis=1)
    w = s.var(axis=0, ddof=1).mean()
    b = n * s.mean(axis=0).var(ddof=1)
    return float(np.sqrt(((n - 1) / n * w + b / n) / w))


def test_c1_fp64_parity_where_the_link_saturates():
    d = W.logistic_small()
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    ref = O.GlmData(O.LOGIT, d["X"], d["y"], prior=d["prior"])
    th = W.random_thetas(m.D, 20, seed=1001)
    eta = d["X"] @ th[:, 1:].T + th[:, 0]
    assert (np.abs(eta) > 20).sum() > 100  # the saturated regime is exercised
    worst_lp = worst_g = 0.0
    for t in th:
        lp, g = m.log_prob_grad(t)
        st, lp_r, g_r = ref.lpgrad(t, mode=1)
        assert st == 0
        worst_lp = max(worst_lp, abs(lp - lp_r) / abs(lp_r))
        worst_g = max(worst_g, np.linalg.norm(g - g_r) / np.linalg.norm(g_r))
    _say(f"C1 FP64: max rel lp err {worst_lp:.2e}, grad {worst_g:.2e}")
    assert worst_lp <= 1e-12 and worst_g <= 1e-12
    m.close()


def test_c1_device_chains_match_oracle_sampler():
    import threading
    d = W.logistic_small()
    m = S.GlmModel(S.LOGIT, d["X"], d["y"], prior=d["prior"])
    cfg = dict(num_warmup=1000, num_samples=1000, max_depth=10)
    res = S.hmc_nuts_diag_e_adapt_batch(m, 4, random_seed=1234, chain=1, **cfg)
    assert np.all(res["status"] == 0), res["status"]
    dev = np.transpose(res["draws"][:, :, 7:], (1, 0, 2))
    ref = O.GlmData(O.LOGIT, d["X"], d["y"], prior=d["prior"])
    f = O.model_fn("orc_glm_model_cb")
    ctx = ref.ctx(mode=0)
    out = [None] * 4

    def run(c):
        out[c] = O.hmc_nuts_diag_e_adapt(f, ctx, m.D, seed=4321, chain=c + 1, **cfg)
    ths = [threading.Thread(target=run, args=(c,)) for c in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert all(o[0] == 0 for o in out)
    cpu = np.stack([o[1][:, 7:] for o in out], 1)
    bad = []
    for p in range(m.D):
        se = np.hypot(analyze.mcse_mean(dev[:, :, p]), analyze.mcse_mean(cpu[:, :, p]))
        if abs(dev[:, :, p].mean() - cpu[:, :, p].mean()) >= 4 * se:
            bad.append(("mean", p))
        se_sd = np.hypot(analyze.mcse_sd(dev[:, :, p]), analyze.mcse_sd(cpu[:, :, p]))
        if abs(dev[:, :, p].std() - cpu[:, :, p].std()) >= 4 * se_sd:
            bad.append(("sd", p))
    _say(f"C1 4 chains: min bulk ESS device {analyze.min_bulk_ess(dev):.0f}, "
         f"oracle {analyze.min_bulk_ess(cpu):.0f}; outside 4 MCSE: {bad}")
    assert not bad
    m.close()


@pytest.mark.parametrize("cfg", ["C2", "C3"])
def test_mixed_precision_error_at_full_size(cfg):
    if cfg == "C2":
        d, fam = W.logistic_large(), S.LOGIT
    else:
        d, fam = W.normal_batched(), S.NORMAL
    m = S.GlmModel(fam, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)
    ref = O.GlmData(fam, d["X"], d["y"], prior=d["prior"])
    th = W.random_thetas(m.D, 256, seed=2002)
    res = {}
    for scale in (0.5, 1.0):
        full = th * scale
        # 256 chains evaluated together; 8 of them checked against the oracle
        lp, g = m.log_prob_grad_batch(full)
        e_lp = e_g = 0.0
        for c in range(0, 256, 32):
            st, lp_r, g_r = ref.lpgrad(full[c], mode=1)
            assert st == 0
            e_lp = max(e_lp, abs(lp[c] - lp_r) / abs(lp_r))
            e_g = max(e_g, float(np.linalg.norm(g[c] - g_r) / np.linalg.norm(g_r)))
        res[scale] = (e_lp, e_g)
        _say(f"{cfg} BF16 hi+lo, theta ~ N(0, {scale}^2): max rel lp err {e_lp:.2e}, "
             f"grad {e_g:.2e}")
    assert res[0.5][1] <= 1e-6 and res[1.0][1] <= 3e-6
    assert res[0.5][0] <= 1e-7 and res[1.0][0] <= 1e-7
    m.close()


def test_c3_sampler_at_full_size():
    d = W.normal_batched()
    mb = S.GlmModel(S.NORMAL, d["X"], d["y"], prior=d["prior"], precision=S.BF16X2)
    cfg = dict(num_warmup=300, num_samples=300, random_seed=303, chain=1)
    rb = S.hmc_nuts_diag_e_adapt_batch(mb, 256, grad_precision=S.BF16X2, **cfg)
    mb.close()
    assert np.all(rb["status"] == 0), rb["status"]
    bf = np.transpose(rb["draws"][:, :, 7:], (1, 0, 2))  # iterations x chains x params
    rhat = max(_split_rhat(bf[:, :, p]) for p in range(bf.shape[2]))
    m64 = S.GlmModel(S.NORMAL, d["X"], d["y"], prior=d["prior"])
    r64 = S.hmc_nuts_diag_e_adapt_batch(m64, 16, grad_precision=S.FP64,
                                        **{**cfg, "chain": 1001})
    m64.close()
    assert np.all(r64["status"] == 0), r64["status"]
    f64 = np.transpose(r64["draws"][:, :, 7:], (1, 0, 2))
    bad = []
    for p in range(bf.shape[2]):
        se = np.hypot(analyze.mcse_mean(bf[:, :, p]), analyze.mcse_mean(f64[:, :, p]))
        if abs(bf[:, :, p].mean() - f64[:, :, p].mean()) >= 4 * se:
            bad.append(p)
    _say(f"C3 256 chains BF16: max split R-hat {rhat:.4f}, min bulk ESS "
         f"{analyze.min_bulk_ess(bf):.0f}; pooled means outside 4 MCSE of FP64: {bad}")
    assert rhat < 1.01
    assert not bad


def test_c4_fp32_gradient_and_sampler_at_full_size():
    d = W.hier_logistic()
    m = S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                   precision=S.FP32)
    ref = O.GlmData(O.HIER, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"])
    th = W.random_thetas(m.D, 64, seed=4004) * 0.5
    lp, g = m.log_prob_grad_batch(th)
    e_lp = e_g = 0.0
    for c in range(0, 64, 16):
        st, lp_r, g_r = ref.lpgrad(th[c], mode=1)
        assert st == 0
        e_lp = max(e_lp, abs(lp[c] - lp_r) / abs(lp_r))
        e_g = max(e_g, float(np.linalg.norm(g[c] - g_r) / np.linalg.norm(g_r)))
    _say(f"C4 FP32 batched (N={d['X'].shape[0]}, D={m.D}): max rel lp err {e_lp:.2e}, "
         f"grad {e_g:.2e}")
    assert e_lp <= 1e-6 and e_g <= 1e-5
    # Stan's default warm-up (1000 iterations, max_depth 10): after 200
    # warm-up iterations at max_depth 8, chains started far out in the
    # 21,042-dimensional space were still on their way in (step sizes 100x
    # those of converged chains, lp thousands below the mode)
    b = S.NutsBatch(m, 64, grad_precision=S.FP32, random_seed=404, chain=1, num_warmup=1000,
                    num_samples=200, max_depth=10)
    done = 0
    while done < 64:
        steps, done = b.run(20000)
        _say(f"C4 sampler: {steps} steps, {done} of 64 chains done")
        if steps == 0:
            break
    res = b.results()
    b.close()
    assert np.all(res["status"] == 0), res["status"]
    K = d["X"].shape[1]
    dr = res["draws"]
    q = np.transpose(dr[:, :, 7:], (1, 0, 2))
    glob = q[:, :, : 1 + K + K + 1]  # alpha, beta, tau
    rhat = max(_split_rhat(glob[:, :, p]) for p in range(glob.shape[2]))
    ess = analyze.min_bulk_ess(glob)
    # every chain in the typical set: its mean lp within 4 within-chain sds
    # of the median chain's, step size within 2x of the median, no more
    # than 1% divergent transitions.  (Chains stuck outside it, as after a
    # 200-iteration warm-up, were thousands of lp units below the mode with
    # split R-hat 7.)  With 200 draws the per-chain ESS of these parameters
    # is near 10, and split R-hat then sits near sqrt(1 + 1/10) = 1.05 for
    # chains that agree; it is bounded at 1.2.
    lp = dr[:, :, 0]
    lp_sd = lp.std(axis=1).mean()
    lp_far = np.abs(lp.mean(axis=1) - np.median(lp.mean(axis=1))) / lp_sd
    eps = res["stepsize"]
    _say(f"C4 64 chains FP32: max split R-hat of the global parameters {rhat:.4f}, "
         f"min bulk ESS {ess:.0f}; max |lp - median| {lp_far.max():.2f} sd; step sizes "
         f"{eps.min():.3g}..{eps.max():.3g}; divergent {dr[:, :, 5].mean():.4f}")
    assert lp_far.max() < 4.0
    assert eps.max() < 2 * np.median(eps) and eps.min() > 0.5 * np.median(eps)
    assert dr[:, :, 5].mean() < 0.01
    assert rhat < 1.2
    m.close()


def test_c5_largest_shapes_same_alone_and_in_mix():
    """The serving stream's largest shapes (bench seed 7: K = 468 with 64
    chains; N = 8e5 rows): each job's draws are bitwise the same alone and
    sharing the device."""
    big_k = W.logistic(48_176, 468, 1.0 / np.sqrt(468), 0.5, seed=701)
    big_n = W.normal(800_000, 24, seed=702)
    jobs = {
        "k468": dict(family=S.LOGIT, X=big_k["X"], y=big_k["y"], chains=64,
                     prior=big_k["prior"], grad_precision=S.BF16X2),
        "n8e5": dict(family=S.NORMAL, X=big_n["X"], y=big_n["y"], chains=16,
                     prior=big_n["prior"], grad_precision=S.BF16X2),
    }
    common = dict(num_warmup=100, num_samples=60, random_seed=77, chain=1)

    def submit(srv, spec):
        spec = dict(spec)
        fam, X, y, chains = spec.pop("family"), spec.pop("X"), spec.pop("y"), spec.pop("chains")
        return srv.submit(fam, X, y, chains, **spec, **common)
    alone = {}
    for name, spec in jobs.items():
        srv = S.Server(100 * GB)
        jid = submit(srv, spec)
        assert srv.wait(jid, 300), name
        alone[name] = srv.results(jid)
        srv.close()
        _say(f"C5 alone {name}: {alone[name]['info']['steps']} steps")
    srv = S.Server(100 * GB)
    ids = {name: submit(srv, spec) for name, spec in jobs.items()}
    for name, jid in ids.items():
        assert srv.wait(jid, 300), name
        r = srv.results(jid)
        assert np.all(r["status"] == 0)
        np.testing.assert_array_equal(r["draws"], alone[name]["draws"])
    infos = [srv.info(j) for j in ids.values()]
    assert max(i["t_admit"] for i in infos) < min(i["t_done"] for i in infos)
    srv.close()
