"""The dense_e metric services (hmc_nuts_dense_e_adapt, hmc_nuts_dense_e)
on CPU models: the product's C++ sampler against the oracle draw for draw,
the reference's own dense-metric unit checks restated in C++, the adapted
metric and posterior moments on a correlated Gaussian, configuration errors,
and the Stan CSV metric block."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import stan_amd as S
from stan_amd import analyze
from stan_amd.stan_csv import read_stan_csv

HERE = os.path.dirname(os.path.abspath(__file__))
RHO = 0.9


def _corr_cov(D):
    s = 1.0 + np.arange(D)
    return np.array([[s[i] * s[j] * RHO ** abs(i - j) for j in range(D)] for i in range(D)])


def _oracle_kw(kw):
    okw = dict(kw)
    okw["seed"] = okw.pop("random_seed")
    return okw


@pytest.fixture(scope="module")
def dense_check():
    exe = os.path.join(str(HERE), "..", "_build_dense_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "cpp", "dense_metric_check.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    os.remove(exe)
    return dict(zip(out.split()[0::2], map(float, out.split()[1::2])))


def test_reference_dense_unit_checks(dense_check):
    # covar_adaptation_test.cpp:5-28: EXPECT_EQ on every element
    assert dense_check["covar_adapt_maxdiff"] == 0.0
    # dense_e_metric_test.cpp:52-58: within 5 (Wishart) standard deviations
    assert dense_check["sample_p_sigmas"] < 5.0
    # dense_e_metric_test.cpp:102, 120, 141: EXPECT_NEAR(..., epsilon = 1e-6)
    assert dense_check["dtau_dp_fd"] < 1e-6
    assert dense_check["dphi_dq_fd"] < 1e-6
    # derived_nuts_test.cpp:119, 131
    assert dense_check["criterion_true"] == 1 and dense_check["criterion_false"] == 1


def test_oracle_cholesky_pins():
    A = _corr_cov(6)
    L = O.cholesky(A)
    np.testing.assert_allclose(L, np.linalg.cholesky(A), rtol=1e-13, atol=1e-13)
    assert O.cholesky(np.array([[1.0, 2.0], [2.0, 1.0]])) is None


@pytest.mark.parametrize("chain", [1, 3])
def test_dense_adapt_matches_oracle_bitwise(chain):
    f = O.model_fn("orc_corr_gauss_model_cb")
    kw = dict(random_seed=41, chain=chain, num_warmup=300, num_samples=200, max_depth=8)
    prod = S.hmc_nuts_dense_e_adapt_cb(f, None, 6, **kw)
    rc, draws, eps, minv, ng = O.hmc_nuts_dense_e_adapt(f, None, 6, **_oracle_kw(kw))
    assert prod["rc"] == 0 and rc == 0
    np.testing.assert_array_equal(prod["draws"], draws)
    assert prod["stepsize"] == eps
    np.testing.assert_array_equal(prod["inv_metric"], minv)
    assert prod["n_grad"] == ng


def test_dense_adapt_matches_oracle_bitwise_glm():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((300, 5)) / np.sqrt(5)
    X[:, 1] = 0.8 * X[:, 0] + 0.2 * X[:, 1]  # correlated columns
    y = (rng.random(300) < 0.5).astype(np.int32)
    data = O.GlmData(O.LOGIT, X, y)
    ctx = data.ctx(mode=0)
    f = O.model_fn("orc_glm_model_cb")
    kw = dict(random_seed=1234, chain=1, num_warmup=150, num_samples=150, save_warmup=1,
              num_thin=2)
    prod = S.hmc_nuts_dense_e_adapt_cb(f, C.byref(ctx), data.D, **kw)
    rc, draws, eps, minv, _ = O.hmc_nuts_dense_e_adapt(f, ctx, data.D, **_oracle_kw(kw))
    assert prod["rc"] == 0 and rc == 0
    np.testing.assert_array_equal(prod["draws"], draws)
    np.testing.assert_array_equal(prod["inv_metric"], minv)
    assert prod["draws"].shape == (150, 7 + data.D)


def test_dense_fixed_metric_matches_oracle_bitwise():
    """hmc_nuts_dense_e: no heuristic, no adaptation, the given metric."""
    f = O.model_fn("orc_corr_gauss_model_cb")
    M = _corr_cov(5)
    kw = dict(random_seed=5, chain=2, num_warmup=50, num_samples=300, stepsize=0.7,
              save_warmup=1)
    prod = S.hmc_nuts_dense_e_cb(f, None, 5, inv_metric=M, **kw)
    rc, draws, ng = O.hmc_nuts_dense_e(f, None, 5, init_inv_metric=M, **_oracle_kw(kw))
    assert prod["rc"] == 0 and rc == 0
    np.testing.assert_array_equal(prod["draws"], draws)
    assert np.all(prod["draws"][:, 2] == 0.7)
    np.testing.assert_array_equal(prod["inv_metric"], M)
    assert prod["n_grad"] == ng
    # the non-adaptive diag service agrees with the oracle too
    kw = dict(random_seed=6, chain=1, num_warmup=20, num_samples=100, stepsize=0.4)
    pd = S.hmc_nuts_diag_e_cb(f, None, 5, **kw)
    rc, dd, _ = O.hmc_nuts_diag_e(f, None, 5, **_oracle_kw(kw))
    assert pd["rc"] == 0 and rc == 0
    np.testing.assert_array_equal(pd["draws"], dd)


def test_dense_adaptation_learns_the_covariance():
    """On x ~ N(0, S) the adapted dense metric estimates S, the draws have
    the right moments (4 MCSE), and the dense sampler needs fewer leapfrog
    steps per draw than the diag one (S is far from diagonal)."""
    D = 8
    Sig = _corr_cov(D)
    f = O.model_fn("orc_corr_gauss_model_cb")
    dense, diag = [], []
    for chain in range(1, 5):
        kw = dict(random_seed=2024, chain=chain, num_warmup=1000, num_samples=1000)
        out = S.hmc_nuts_dense_e_adapt_cb(f, None, D, **kw)
        assert out["rc"] == 0
        dense.append(out)
        diag.append(S.hmc_nuts_diag_e_adapt_cb(f, None, D, **kw))
    # the last window saw 500 draws: relative error of the metric ~ sqrt(2/500)
    for out in dense:
        M = out["inv_metric"]
        np.testing.assert_array_equal(M, M)  # finite
        rel = np.linalg.norm(M - Sig) / np.linalg.norm(Sig)
        assert rel < 0.35, rel
    x = np.stack([o["draws"][:, 7:] for o in dense], axis=1)
    for p in range(D):
        se = analyze.mcse_mean(x[:, :, p])
        assert abs(x[:, :, p].mean()) < 4 * se, p
        assert abs(x[:, :, p].var() / Sig[p, p] - 1.0) < 0.2, p
    lf_dense = np.mean([o["draws"][:, 4].mean() for o in dense])
    lf_diag = np.mean([o["draws"][:, 4].mean() for o in diag])
    assert lf_dense < 0.5 * lf_diag, (lf_dense, lf_diag)


def test_dense_config_errors():
    f = O.model_fn("orc_corr_gauss_model_cb")
    kw = dict(num_warmup=10, num_samples=10)
    asym = np.array([[1.0, 0.1], [0.2, 1.0]])
    not_pd = np.array([[1.0, 2.0], [2.0, 1.0]])
    nan = np.array([[1.0, np.nan], [np.nan, 1.0]])
    for M in (asym, not_pd, nan):
        assert S.hmc_nuts_dense_e_adapt_cb(f, None, 2, inv_metric=M, **kw)["rc"] == 78
        assert S.hmc_nuts_dense_e_cb(f, None, 2, inv_metric=M, **kw)["rc"] == 78
        rc = O.hmc_nuts_dense_e_adapt(f, None, 2, init_inv_metric=M, **kw)[0]
        assert rc == 78
    # the reference logs "Inverse Euclidean metric not positive definite."
    # and returns with e.what() of its domain_error logged last
    assert S.last_error() == "Initialization failure"


def test_dense_csv_metric_block(tmp_path):
    f = O.model_fn("orc_corr_gauss_model_cb")
    path = tmp_path / "dense.csv"
    out = S.hmc_nuts_dense_e_adapt_cb_csv(f, None, 3, path, sig_figs=18, random_seed=9,
                                          num_warmup=200, num_samples=50)
    assert out["rc"] == 0
    text = path.read_text()
    assert "metric = dense_e" in text
    assert "# Elements of inverse mass matrix:" in text
    r = read_stan_csv(path)
    assert r["inv_metric"].shape == (3, 3)
    # metric lines use a fresh stream's 6 significant digits, as the
    # reference's dense_e_point::write_metric (dense_e_point.hpp:44-56)
    np.testing.assert_allclose(r["inv_metric"], out["inv_metric"], rtol=5e-6)
    np.testing.assert_allclose(r["samples"], out["draws"], rtol=1e-15)
    assert r["step_size"] == pytest.approx(out["stepsize"], rel=5e-6)
