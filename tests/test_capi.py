"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/stanx.h declares, validates inputs, and its C++ sampler reproduces
the oracle sampler draw for draw on CPU models."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import stan_amd as S
from stan_amd import analyze

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "stanx.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = S.lib()
    declared = _declared_symbols()
    assert len(declared) >= 12
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(S.EXPORTED) == declared


def test_create_rejects_bad_labels_before_touching_the_device():
    X = np.zeros((3, 2))
    y = np.array([0, 2, 1], dtype=np.int32)
    with pytest.raises(S.StanxError) as e:
        S.GlmModel(S.LOGIT, X, y)
    assert e.value.code == S.SX_EINVAL
    assert "must be in {0, 1}" in str(e.value)


def test_create_rejects_nonfinite_data():
    X = np.array([[0.0, np.nan]])
    with pytest.raises(S.StanxError) as e:
        S.GlmModel(S.NORMAL, X, np.zeros(1))
    assert e.value.code == S.SX_EINVAL


def _compare_runs(fn_ptr, ctx, D, **kw):
    prod = S.hmc_nuts_diag_e_adapt_cb(fn_ptr, C.byref(ctx) if ctx is not None else None,
                                      D, **kw)
    okw = dict(kw)
    okw["seed"] = okw.pop("random_seed")
    rc, draws, eps, minv, _ = O.hmc_nuts_diag_e_adapt(fn_ptr, ctx, D, **okw)
    assert prod["rc"] == 0 and rc == 0
    return prod, draws, eps, minv


def test_product_sampler_matches_oracle_bitwise_gauss():
    """The iterative C++ NUTS and the recursive oracle consume the same
    Philox stream and must give identical draws (trajectories, multinomial
    choices, adaptation) on the same model."""
    f = O.model_fn("orc_gauss_model_cb")
    for chain in (1, 2):
        prod, draws, eps, minv = _compare_runs(f, None, 5, random_seed=99, chain=chain,
                                               num_warmup=300, num_samples=300,
                                               max_depth=8)
        np.testing.assert_array_equal(prod["draws"], draws)
        assert prod["stepsize"] == eps
        np.testing.assert_array_equal(prod["inv_metric"], minv)


def test_product_sampler_matches_oracle_bitwise_glm():
    rng = np.random.default_rng(5)
    X = rng.standard_normal((400, 6)) / np.sqrt(6)
    y = (rng.random(400) < 0.5).astype(np.int32)
    data = O.GlmData(O.LOGIT, X, y)
    ctx = data.ctx(mode=0)
    f = O.model_fn("orc_glm_model_cb")
    prod, draws, eps, minv = _compare_runs(f, ctx, data.D, random_seed=1234, chain=1,
                                           num_warmup=200, num_samples=200, save_warmup=1,
                                           num_thin=3)
    np.testing.assert_array_equal(prod["draws"], draws)
    assert prod["draws"].shape == ((200 + 200 + 2) // 3, 7 + data.D)


def test_product_sampler_funnel_divergences_match_oracle():
    f = O.model_fn("orc_funnel_model_cb")
    prod, draws, _, _ = _compare_runs(f, None, 11, random_seed=3, chain=1, num_warmup=150,
                                      num_samples=150)
    np.testing.assert_array_equal(prod["draws"], draws)
    assert draws[:, 5].sum() >= 0  # divergent__ column present


def test_service_config_error_codes():
    f = O.model_fn("orc_gauss_model_cb")
    bad_metric = np.array([1.0, -1.0])
    out = S.hmc_nuts_diag_e_adapt_cb(f, None, 2, inv_metric=bad_metric, num_warmup=10,
                                     num_samples=10)
    assert out["rc"] == 78


def test_non_adaptive_service_fixed_stepsize_and_moments():
    """hmc_nuts_diag_e: the same transition with the step size and metric
    held fixed (no heuristic, no adaptation); gradient accounting of
    base_nuts.hpp:85 (one per transition) plus the initialisation."""
    f = O.model_fn("orc_gauss_model_cb")
    draws = []
    for chain in range(1, 5):
        out = S.hmc_nuts_diag_e_cb(f, None, 3, random_seed=77, chain=chain, num_warmup=100,
                                   num_samples=1000, save_warmup=1, stepsize=0.9)
        assert out["rc"] == 0
        d = out["draws"]
        assert d.shape == (1100, 10)
        assert np.all(d[:, 2] == 0.9)
        assert out["stepsize"] == 0.9
        np.testing.assert_array_equal(out["inv_metric"], np.ones(3))
        assert out["n_grad"] == int(d[:, 4].sum()) + d.shape[0] + 1
        draws.append(d[100:, 7:])
    x = np.stack(draws, axis=1)
    for p in range(3):
        assert abs(x[:, :, p].mean()) < 4 * analyze.mcse_mean(x[:, :, p])
        assert abs(x[:, :, p].var() - 1.0) < 0.15
    # jitter: step sizes uniform in eps (1 +- j)
    out = S.hmc_nuts_diag_e_cb(f, None, 3, random_seed=5, num_warmup=0, num_samples=200,
                               stepsize=0.5, stepsize_jitter=0.4)
    eps = out["draws"][:, 2]
    assert eps.min() >= 0.3 and eps.max() <= 0.7 and eps.std() > 0.05
    # a bad metric is a configuration error, as in the adaptive service
    out = S.hmc_nuts_diag_e_cb(f, None, 2, inv_metric=np.array([1.0, 0.0]), num_samples=5)
    assert out["rc"] == 78
