"""Hierarchical logistic model (ragged groups, non-centred varying
intercepts and slopes): segmented GPU passes against the oracle.

Bars:
  FP64 (single chain and batched): log density and gradient within 1e-12
  of the compensated oracle, relative (normwise for the gradient); a chain's
  batched result is bitwise its single evaluation.
  FP32 batched (X in FP32; FP32 matrix-core products and FP32 sums within
  a chunk of at most 256 rows by default (128 .. 1024 when autotuned),
  FP32 row terms per 32-row tile added in FP64; FP64 sums over chunks and
  groups), random test points
  theta ~ N(0, 0.5^2):
  relative error of the log density <= LP_TOL_FP32 and normwise relative
  error of the gradient <= GRAD_TOL_FP32; the measured values are printed.
"""
import sys

import numpy as np
import pytest

import oracle as O
import stan_amd as S
from stan_amd import analyze
from stan_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-12
LP_TOL_FP32 = 1e-5
GRAD_TOL_FP32 = 1e-4


def _nanmax(a, b):
    """max that keeps a NaN (the builtin max(a, nan) returns a)."""
    return a if a != a else (b if b != b else max(a, b))


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def _data(J, K, mean_size, seed, empty=()):
    d = W.hier_logistic(J=J, K=K, seed=seed, mean_log_size=np.log(mean_size), sd_log=1.0)
    if empty:
        # drop the rows of some groups: empty groups keep their z
        sizes = np.diff(d["offsets"])
        keep = np.ones(d["X"].shape[0], bool)
        for j in empty:
            keep[d["offsets"][j]:d["offsets"][j + 1]] = False
            sizes[j] = 0
        d["X"], d["y"] = d["X"][keep], d["y"][keep]
        d["offsets"] = np.concatenate([[0], np.cumsum(sizes)])
    return d


def _models(d, precision=S.FP64):
    m = S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                   precision=precision)
    ref = O.GlmData(O.HIER, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"])
    return m, ref


@pytest.mark.parametrize("J,K,size,empty", [(1, 1, 10, ()), (3, 0, 40, ()), (5, 3, 60, (1,)),
                                            (40, 4, 300, (0, 39)), (17, 8, 900, ()),
                                            (50, 20, 400, (7,)), (9, 24, 250, ()),
                                            (12, 32, 500, (3, 4))])
def test_hier_fp64_parity(J, K, size, empty):
    d = _data(J, K, size, seed=J * 100 + K, empty=empty)
    _say(f"hier fp64 J={J} K={K} N={d['X'].shape[0]} empty={empty}")
    m, ref = _models(d)
    assert m.D == 2 * K + 2 + J * (K + 1)
    for th in W.random_thetas(m.D, 3, seed=K + J) * 0.5:
        for propto, jac in [(True, True), (False, False), (False, True)]:
            lp, g = m.log_prob_grad(th, propto=propto, jacobian=jac)
            st, lp_r, g_r = ref.lpgrad(th, mode=1, propto=propto, jacobian=jac)
            assert st == 0
            assert abs(lp - lp_r) <= TOL * max(1.0, abs(lp_r)), (lp, lp_r)
            err = np.linalg.norm(g - g_r)
            assert err <= TOL * max(1.0, np.linalg.norm(g_r)), (err, np.linalg.norm(g_r))
    m.close()


def test_hier_fp64_batch_bitwise_single():
    d = _data(30, 6, 200, seed=11)
    m, ref = _models(d)
    th = W.random_thetas(m.D, 9, seed=3) * 0.5
    lp, g = m.log_prob_grad_batch(th)
    for c in range(9):
        lp1, g1 = m.log_prob_grad(th[c])
        assert lp[c] == lp1
        np.testing.assert_array_equal(g[c], g1)
    m.close()


@pytest.mark.parametrize("J,K,size,C", [(5, 3, 60, 1), (40, 4, 300, 7), (50, 20, 400, 64),
                                        (12, 32, 500, 130), (3, 0, 40, 65)])
def test_hier_fp32_batch_against_oracle(J, K, size, C):
    d = _data(J, K, size, seed=J + K + C)
    m, ref = _models(d, precision=S.FP32)
    th = W.random_thetas(m.D, C, seed=C + K) * 0.5
    lp, g = m.log_prob_grad_batch(th)
    worst_lp = worst_g = 0.0
    for c in range(C):
        st, lp_r, g_r = ref.lpgrad(th[c], mode=1)
        assert st == 0
        worst_lp = _nanmax(worst_lp, abs(lp[c] - lp_r) / abs(lp_r))
        worst_g = _nanmax(worst_g, np.linalg.norm(g[c] - g_r) / np.linalg.norm(g_r))
    _say(f"hier fp32 J={J} K={K} C={C}: max rel lp err {worst_lp:.2e}, "
         f"max normwise grad err {worst_g:.2e}")
    assert worst_lp < LP_TOL_FP32 and worst_g < GRAD_TOL_FP32
    # single evaluations of the same context stay FP64
    lp64, g64 = m.log_prob_grad(th[0])
    st, lp_r, g_r = ref.lpgrad(th[0], mode=1)
    assert abs(lp64 - lp_r) <= TOL * abs(lp_r)
    m.close()


@pytest.mark.parametrize("chunk", [32, 128, 512, 1024])
def test_hier_chunk_sizes_against_oracle(chunk):
    """Every chunk size the autotuner may pick (and the smallest allowed)
    gives the oracle's values: FP64 single chain to 1e-12, FP32 batch to
    its bars.  Groups of up to a few thousand rows span many chunks."""
    d = _data(20, 6, 1500, seed=chunk)
    _say(f"hier chunk={chunk} N={d['X'].shape[0]}")
    ref = O.GlmData(O.HIER, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"])
    th = W.random_thetas(2 * 6 + 2 + 20 * 7, 40, seed=chunk + 1) * 0.5
    for precision in (S.FP64, S.FP32):
        m = S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                       precision=precision, tuning=dict(hier_chunk=chunk))
        if precision == S.FP64:
            for c in range(3):
                lp, g = m.log_prob_grad(th[c])
                st, lp_r, g_r = ref.lpgrad(th[c], mode=1)
                assert st == 0
                assert abs(lp - lp_r) <= TOL * max(1.0, abs(lp_r)), (lp, lp_r)
                assert np.linalg.norm(g - g_r) <= TOL * max(1.0, np.linalg.norm(g_r))
        else:
            lp, g = m.log_prob_grad_batch(th)
            worst_lp = worst_g = 0.0
            for c in range(th.shape[0]):
                st, lp_r, g_r = ref.lpgrad(th[c], mode=1)
                worst_lp = _nanmax(worst_lp, abs(lp[c] - lp_r) / abs(lp_r))
                worst_g = _nanmax(worst_g, np.linalg.norm(g[c] - g_r) / np.linalg.norm(g_r))
            _say(f"  fp32: max rel lp err {worst_lp:.2e}, max normwise grad err {worst_g:.2e}")
            assert worst_lp < LP_TOL_FP32 and worst_g < GRAD_TOL_FP32
        m.close()


def test_hier_autotuned_contexts_are_bitwise_equal():
    """The chunk size sets the order of the FP64 sums; the autotuner's first
    decision per device and shape is reused, so two autotuned contexts of
    one data set give bitwise-equal results (FP64 single chain and FP32
    batch)."""
    d = _data(150, 8, 4000, seed=77)
    N = d["X"].shape[0]
    assert N * 8 >= 1 << 22
    _say(f"hier autotune N={N}")
    for precision in (S.FP64, S.FP32):
        ms = [S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                         precision=precision, tuning=dict(autotune=1)) for _ in range(2)]
        log0, log1 = ms[0].tuning_log(), ms[1].tuning_log()
        _say(log0.strip())
        assert "hierarchical chunk" in log0 and "hierarchical chunk cached" in log1, (log0, log1)
        th = W.random_thetas(ms[0].D, 8, seed=5) * 0.5
        if precision == S.FP64:
            a, b = ms[0].log_prob_grad(th[0]), ms[1].log_prob_grad(th[0])
        else:
            a, b = ms[0].log_prob_grad_batch(th), ms[1].log_prob_grad_batch(th)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        for m in ms:
            m.close()


@pytest.mark.parametrize("J,K,size,C", [(5, 3, 60, 1), (40, 4, 300, 7), (50, 20, 400, 64),
                                        (12, 31, 500, 130), (3, 0, 40, 65), (9, 15, 700, 33),
                                        (8, 16, 300, 32)])
def test_hier_fp32_split_products_against_oracle(J, K, size, C):
    """FP32 batches with the products formed on BF16 MFMA from exact
    three-term splits: the same bars as the FP32 MFMA products."""
    d = _data(J, K, size, seed=J + K + C + 1)
    m = S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                   precision=S.FP32, tuning=dict(hier_kernel=2))
    ref = O.GlmData(O.HIER, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"])
    th = W.random_thetas(m.D, C, seed=C + K + 1) * 0.5
    lp, g = m.log_prob_grad_batch(th)
    worst_lp = worst_g = 0.0
    for c in range(C):
        st, lp_r, g_r = ref.lpgrad(th[c], mode=1)
        assert st == 0
        worst_lp = _nanmax(worst_lp, abs(lp[c] - lp_r) / abs(lp_r))
        worst_g = _nanmax(worst_g, np.linalg.norm(g[c] - g_r) / np.linalg.norm(g_r))
    _say(f"hier split products J={J} K={K} C={C}: max rel lp err {worst_lp:.2e}, "
         f"max normwise grad err {worst_g:.2e}")
    assert worst_lp < LP_TOL_FP32 and worst_g < GRAD_TOL_FP32
    # a chain's result does not depend on the batch
    sel = np.arange(C)[::3]
    lp_s, g_s = m.log_prob_grad_batch(th[sel])
    np.testing.assert_array_equal(lp_s, lp[sel])
    np.testing.assert_array_equal(g_s, g[sel])
    m.close()


def test_hier_split_products_domain_error_per_chain():
    d = _data(6, 3, 50, seed=9)
    m = S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                   precision=S.FP32, tuning=dict(hier_kernel=2))
    th = W.random_thetas(m.D, 4, seed=2) * 0.5
    th[2, 1 + 3] = 800.0  # tau_0 = exp(800) overflows: eta is not finite
    lp, g = m.log_prob_grad_batch(th)
    assert np.isneginf(lp[2]) and np.all(np.isfinite(lp[[0, 1, 3]]))
    m.close()


def test_hier_fp32_chain_independent_of_batch():
    d = _data(25, 5, 150, seed=5)
    m, _ = _models(d, precision=S.FP32)
    th = W.random_thetas(m.D, 100, seed=6) * 0.5
    lp_all, g_all = m.log_prob_grad_batch(th)
    perm = np.random.Generator(np.random.PCG64(1)).permutation(100)[:37]
    lp_p, g_p = m.log_prob_grad_batch(th[perm])
    np.testing.assert_array_equal(lp_p, lp_all[perm])
    np.testing.assert_array_equal(g_p, g_all[perm])
    m.close()


@pytest.mark.parametrize("precision", [S.FP64, S.FP32])
def test_hier_domain_error_per_chain(precision):
    d = _data(6, 3, 50, seed=9)
    m, _ = _models(d, precision=precision)
    th = W.random_thetas(m.D, 4, seed=2) * 0.5
    th[2, 1 + 3] = 800.0  # tau_0 = exp(800) overflows: eta is not finite
    lp, g = m.log_prob_grad_batch(th)
    assert np.isneginf(lp[2]) and np.all(np.isfinite(lp[[0, 1, 3]]))
    with pytest.raises(S.DomainError):
        m.log_prob_grad(th[2])
    m.close()


def test_hier_create_validation():
    d = _data(4, 2, 20, seed=1)
    X, y, off = d["X"], d["y"], d["offsets"]
    bad = off.copy()
    bad[-1] += 1
    with pytest.raises(S.StanxError):
        S.GlmModel(S.HIER_LOGIT, X, y, offsets=bad)
    bad = off.copy()
    bad[1], bad[2] = bad[2], bad[1]
    with pytest.raises(S.StanxError):
        S.GlmModel(S.HIER_LOGIT, X, y, offsets=bad)
    with pytest.raises(S.StanxError):
        S.GlmModel(S.HIER_LOGIT, X, y)
    with pytest.raises(S.StanxError):
        S.GlmModel(S.HIER_LOGIT, X, y, offsets=off, precision=S.BF16X2)
    wide = np.zeros((X.shape[0], 33))
    with pytest.raises(S.StanxError):
        S.GlmModel(S.HIER_LOGIT, wide, y, offsets=off)


def _oracle_posterior(d, D, chains, warmup, samples, seed):
    ref = O.GlmData(O.HIER, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"])
    f = O.model_fn("orc_glm_model_cb")
    ctx = ref.ctx(mode=0)
    out = []
    for chain in range(1, chains + 1):
        rc, dr, _, _, _ = O.hmc_nuts_diag_e_adapt(f, ctx, D, seed=seed, chain=chain,
                                                  num_warmup=warmup, num_samples=samples)
        assert rc == 0
        out.append(dr[:, 7:])
    return np.stack(out, 1)  # iterations x chains x params (unconstrained)


@pytest.mark.parametrize("precision", [S.FP64, S.FP32])
def test_hier_batch_posterior_matches_oracle_sampler(precision):
    d = _data(6, 2, 80, seed=21)
    m, _ = _models(d, precision=precision)
    K = 2
    res = S.hmc_nuts_diag_e_adapt_batch(m, 8, grad_precision=precision, random_seed=17,
                                        chain=1, num_warmup=500, num_samples=1000)
    assert np.all(res["status"] == 0), res["status"]
    dev = np.transpose(res["draws"][:, :, 7:], (1, 0, 2)).copy()
    # draws carry tau = exp(v); compare on the unconstrained scale
    dev[:, :, 1 + K:2 + 2 * K] = np.log(dev[:, :, 1 + K:2 + 2 * K])
    cpu = _oracle_posterior(d, m.D, 4, 500, 1000, seed=4711)
    bad = []
    for p in range(m.D):
        se = np.hypot(analyze.mcse_mean(dev[:, :, p]), analyze.mcse_mean(cpu[:, :, p]))
        if abs(dev[:, :, p].mean() - cpu[:, :, p].mean()) >= 4 * se:
            bad.append(p)
    _say(f"hier batch precision={precision}: min bulk ESS {analyze.min_bulk_ess(dev):.0f}, "
         f"steps {res['steps']}, params outside 4 MCSE: {bad}")
    assert not bad
    m.close()


def test_hier_single_chain_service():
    d = _data(5, 2, 60, seed=33)
    m, _ = _models(d)
    res = S.hmc_nuts_diag_e_adapt(m, random_seed=5, chain=1, num_warmup=200, num_samples=200)
    assert res["rc"] == 0
    dr = res["draws"]
    assert dr.shape == (200, 7 + m.D)
    # write_array: tau columns are exp(v) > 0
    assert np.all(dr[:, 7 + 1 + 2:7 + 2 + 2 * 2] > 0)
    m.close()


@pytest.mark.parametrize("J,K", [(250, 8), (1000, 8)])
def test_hier_batch_large_d_follows_host_service(J, K):
    """D = 2268 and 9018: the device sampler runs 4 and 8 waves per chain
    (block reductions), the host service one thread; same stream, same
    phases, so the draws agree until last-bit differences of the dot
    products grow (see the batched sampler tests)."""
    d = _data(J, K, 40, seed=J + K)
    m, _ = _models(d)
    _say(f"hier large-D batch vs host J={J} K={K} D={m.D} N={d['X'].shape[0]}")
    cfg = dict(random_seed=99, chain=1, num_warmup=40, num_samples=20, save_warmup=1,
               max_depth=5)
    res = S.hmc_nuts_diag_e_adapt_batch(m, 2, **cfg)
    assert np.all(res["status"] == 0), res["status"]
    for c in range(2):
        host = S.hmc_nuts_diag_e_adapt(m, **{**cfg, "chain": 1 + c})
        assert host["rc"] == 0
        dev, hd = res["draws"][c], host["draws"]
        assert dev.shape == hd.shape
        np.testing.assert_allclose(dev[:10], hd[:10], rtol=1e-5, atol=1e-8)
        discrete = np.all(dev[:, 3:6] == hd[:, 3:6], axis=1)
        _say(f"chain {c}: discrete columns equal in {discrete.mean():.3f} of rows")
        assert discrete[:20].all()
    m.close()


def test_multi_part_chains_match_single_part():
    """D > 8192: each chain spreads over several workgroups whose block
    totals are combined across parts.  Only the summation order of the
    reductions changes, so the first draws agree with the one-workgroup
    run to rounding, and every chain ends normally."""
    import os
    d = W.hier_logistic(J=500, K=20, seed=31, mean_log_size=np.log(80))
    m = S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                   precision=S.FP32)
    assert m.D > 8192
    runs = {}
    for parts in ("4", "1"):
        os.environ["STANX_NUTS_PARTS"] = parts
        try:
            b = S.NutsBatch(m, 16, grad_precision=S.FP32, random_seed=5, chain=1,
                            num_warmup=20, num_samples=10, max_depth=5, save_warmup=1)
        finally:
            del os.environ["STANX_NUTS_PARTS"]
        while True:
            steps, done = b.run(max_steps=512)
            if done >= 16 or steps == 0:
                break
        runs[parts] = b.results()
        b.close()
    a, s1 = runs["4"], runs["1"]
    assert np.all(a["status"] == 0) and np.all(s1["status"] == 0)
    np.testing.assert_allclose(a["draws"][:, :3], s1["draws"][:, :3], rtol=1e-6, atol=1e-9)
    m.close()


def test_phase_budget_yields_leave_draws_unchanged():
    """A chain that reaches its per-launch phase budget stops without a
    gradient request and continues in the next step.  The arithmetic and
    the random stream are those of an unbounded run, so the draws (and
    every diagnostic column) are bitwise the same with a budget of 2 turns
    as with none, while the run takes more steps."""
    import os
    d = W.hier_logistic(J=500, K=20, seed=31, mean_log_size=np.log(80))
    m = S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                   precision=S.FP32)
    runs, nsteps = {}, {}
    for budget in ("2", "4096"):
        os.environ["STANX_NUTS_PHASE_BUDGET"] = budget
        try:
            b = S.NutsBatch(m, 16, grad_precision=S.FP32, random_seed=5, chain=1,
                            num_warmup=20, num_samples=10, max_depth=5, save_warmup=1)
        finally:
            del os.environ["STANX_NUTS_PHASE_BUDGET"]
        n = 0
        while True:
            steps, done = b.run(max_steps=512)
            n += steps
            if done >= 16 or steps == 0:
                break
        runs[budget], nsteps[budget] = b.results(), n
        b.close()
    _say(f"steps with budget 2: {nsteps['2']}, unbounded: {nsteps['4096']}")
    a, u = runs["2"], runs["4096"]
    assert np.all(a["status"] == 0) and np.all(u["status"] == 0)
    assert np.array_equal(a["draws"], u["draws"])
    assert nsteps["2"] > nsteps["4096"]
    m.close()


def test_failed_exchange_ends_every_multi_part_chain():
    """A cross-part exchange that times out ends its chain with status 70 in
    every part.  Forced from the host (STANX_NUTS_XFAIL_TEST: every wait
    expects a tag no part writes and gives up after a short bounded wait), so
    every launch still runs to completion.  A batch created afterwards, with
    the hook off, runs normally."""
    import os
    d = W.hier_logistic(J=500, K=20, seed=31, mean_log_size=np.log(80))
    m = S.GlmModel(S.HIER_LOGIT, d["X"], d["y"], offsets=d["offsets"], prior=d["prior"],
                   precision=S.FP32)
    assert m.D > 8192
    _say("forced exchange failure, 16 chains x 4 parts")
    os.environ["STANX_NUTS_PARTS"] = "4"
    os.environ["STANX_NUTS_XFAIL_TEST"] = "1"
    try:
        b = S.NutsBatch(m, 16, grad_precision=S.FP32, random_seed=5, chain=1,
                        num_warmup=20, num_samples=10, max_depth=5)
    finally:
        del os.environ["STANX_NUTS_XFAIL_TEST"]
    total = 0
    while True:
        steps, done = b.run(max_steps=64)
        total += steps
        if done >= 16 or steps == 0 or total > 4096:
            break
    res = b.results()
    b.close()
    assert np.all(res["status"] == 70), res["status"]
    assert total < 64, total  # each chain stops at its first exchange
    _say("hook off: the same batch runs normally")
    try:
        b = S.NutsBatch(m, 16, grad_precision=S.FP32, random_seed=5, chain=1,
                        num_warmup=20, num_samples=10, max_depth=5)
    finally:
        del os.environ["STANX_NUTS_PARTS"]
    while True:
        steps, done = b.run(max_steps=512)
        if done >= 16 or steps == 0:
            break
    res = b.results()
    b.close()
    assert np.all(res["status"] == 0), res["status"]
    m.close()
