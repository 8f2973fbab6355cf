"""Stan JSON data files (src/stan/io/json/json_data.hpp) and an end-to-end
posterior check on the reference's own data set.

The posterior check uses src/test/unit/services/pathfinder/normal_glm_test.json
(copied to tests/golden) with the model of
src/test/test-models/good/services/normal_glm.stan: normal_id_glm(Y | X,
Intercept, b, sigma), Intercept ~ normal(0, 3), b ~ normal(0, 3),
sigma ~ normal(1, 2) on sigma > 0.  The reference records posterior means
for it (src/test/unit/services/pathfinder/util.hpp:495-503) and tests its
Pathfinder means against them at 0.01 (normal_glm_test.cpp:129-132); the
same tolerance is used here for NUTS draws."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle as O
import stan_amd as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
GLM_JSON = os.path.join(GOLDEN, "normal_glm_test.json")
BRMS_PRIOR = (3.0, 3.0, 2.0, 0.0, 1.0)  # alpha_sd, beta_sd, sigma_sd, tau_sd, sigma_mu


def test_json_var_matches_python_json():
    d = json.load(open(GLM_JSON))
    np.testing.assert_array_equal(S.json_var(GLM_JSON, "X"), np.array(d["X"]))
    np.testing.assert_array_equal(S.json_var(GLM_JSON, "Y"), np.array(d["Y"]))
    assert S.json_var(GLM_JSON, "N") == 5000 and S.json_var(GLM_JSON, "K") == 5


def test_json_edge_cases(tmp_path):
    p = tmp_path / "d.json"
    p.write_text('{"a": [[[1, 2], [3, 4], [5, 6]], [[7, 8], [9, 10], [11, 12]]],'
                 ' "e": [], "s": -2.5e-3, "n": ["NaN", "Inf", "-Inf", 1]}')
    a = S.json_var(p, "a")
    np.testing.assert_array_equal(a, np.arange(1, 13).reshape(2, 3, 2))
    assert S.json_var(p, "e").shape == (0,)
    assert S.json_var(p, "s") == -2.5e-3
    n = S.json_var(p, "n")
    assert np.isnan(n[0]) and n[1] == np.inf and n[2] == -np.inf and n[3] == 1
    p.write_text('{"r": [[1, 2], [3]]}')
    with pytest.raises(S.StanxError):
        S.json_var(p, "r")
    with pytest.raises(S.StanxError):
        S.json_var(GLM_JSON, "missing")


def test_sigma_prior_location_gradient():
    rng = np.random.Generator(np.random.PCG64(3))
    X = rng.standard_normal((40, 3))
    y = X @ np.array([1.0, -2.0, 0.5]) + 0.3 + rng.standard_normal(40)
    ref = O.GlmData(O.NORMAL, X, y, prior=BRMS_PRIOR)
    th = np.array([0.2, 0.9, -1.7, 0.4, np.log(1.3)])
    st, lp, g = ref.lpgrad(th, mode=1, propto=False)
    h = 1e-6
    for i in range(5):
        e = np.zeros(5)
        e[i] = h
        fd = (ref.lpgrad(th + e, mode=1, propto=False)[1] -
              ref.lpgrad(th - e, mode=1, propto=False)[1]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i]))
    # the prior term itself: -0.5 ((sigma - 1) / 2)^2 - log 2 - 0.5 log(2 pi)
    base = O.GlmData(O.NORMAL, X, y, prior=(3.0, 3.0, 0.0, 0.0, 0.0))
    lp0 = base.lpgrad(th, mode=1, propto=False)[1]
    s = 1.3
    want = -0.5 * ((s - 1.0) / 2.0) ** 2 - np.log(2.0) - 0.5 * np.log(2 * np.pi)
    assert abs((lp - lp0) - want) < 1e-12


def test_normal_glm_posterior_matches_reference_record():
    d = json.load(open(GLM_JSON))
    X, Y = np.array(d["X"]), np.array(d["Y"])
    ref = O.GlmData(O.NORMAL, X, Y, prior=BRMS_PRIOR)
    f = O.model_fn("orc_glm_model_cb")
    ctx = ref.ctx(mode=0)
    draws = []
    for chain in (1, 2):
        out = S.hmc_nuts_diag_e_adapt_cb(f, C.byref(ctx), ref.D, random_seed=4321, chain=chain,
                                         num_warmup=500, num_samples=1000)
        assert out["rc"] == 0
        draws.append(out["draws"][:, 7:])
    q = np.concatenate(draws)
    intercept, b, sigma = q[:, 0], q[:, 1:6], np.exp(q[:, 6])
    b_intercept = intercept - b @ X.mean(axis=0)
    want_b = [-3.99891, -2.00917, 0.00428024, 0.990037, 2.98999]
    np.testing.assert_allclose(b.mean(axis=0), want_b, atol=0.01)
    assert abs(intercept.mean() - (-1.02273)) < 0.01
    assert abs(sigma.mean() - 1.01613) < 0.01
    assert abs(b_intercept.mean() - (-0.994856)) < 0.01
    # posterior sds of the record (about 0.0145 for b, 0.0102 for sigma)
    np.testing.assert_allclose(b.std(axis=0), [0.0145116, 0.0145324, 0.0136457, 0.0137487,
                                               0.0144594], rtol=0.2)
