"""Posterior diagnostics (min-ESS, MCSE) against the reference's goldens.

The goldens are the reference's own expectations for its analysis
routines on its own fixture files (copied under tests/golden):
  src/test/unit/analyze/mcmc/ess_basic_test.cpp:43-53 (CmdStan 2.35.0)
  src/test/unit/analyze/mcmc/split_rank_normalized_ess_test.cpp:37-53
  src/test/unit/analyze/mcmc/mcse_test.cpp:37-50
  src/test/unit/analyze/mcmc/autocovariance_test.cpp:8-33 (ar1.csv)
  src/test/unit/analyze/mcmc/split_chains_test.cpp:111-135
"""
import os

import numpy as np
import pytest

from stan_amd import analyze as A

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _bern(col):
    cols = []
    for i in range(1, 5):
        rows = []
        with open(os.path.join(GOLDEN, f"bern{i}.csv")) as f:
            for line in f:
                if line.startswith("#") or line.startswith("lp__") or not line.strip():
                    continue
                rows.append([float(v) for v in line.split(",")])
        cols.append(np.array(rows)[:, col])
    return np.column_stack(cols)  # 1000 x 4


@pytest.fixture(scope="module")
def bern():
    return dict(lp=_bern(0), theta=_bern(7), divergent=_bern(5))


def test_ess_basic(bern):
    assert bern["lp"].shape == (1000, 4)
    assert abs(A.ess(bern["lp"]) - 1335.41366787) < 1e-8
    assert abs(A.ess(bern["theta"]) - 1377.5030282) < 1e-8


def test_split_rank_normalized_ess(bern):
    bulk, tail = A.split_rank_normalized_ess(bern["lp"])
    assert abs(bulk - 1512.7684) < 1e-4 and abs(tail - 1591.9707) < 1e-4
    bulk, tail = A.split_rank_normalized_ess(bern["theta"])
    assert abs(bulk - 1407.5124) < 1e-4 and abs(tail - 1291.7131) < 1e-4
    bulk, tail = A.split_rank_normalized_ess(bern["divergent"])
    assert np.isnan(bulk) and np.isnan(tail)


def test_mcse(bern):
    assert abs(A.mcse_mean(bern["lp"]) - 0.02016) < 1e-4
    assert abs(A.mcse_mean(bern["theta"]) - 0.00323) < 1e-4
    assert abs(A.mcse_sd(bern["lp"]) - 0.03553) < 1e-4
    assert abs(A.mcse_sd(bern["theta"]) - 0.00216) < 1e-4
    assert np.isnan(A.mcse_mean(bern["divergent"]))
    th = bern["theta"].copy()
    th[0, 0] = np.inf
    assert np.isnan(A.mcse_mean(th)) and np.isnan(A.mcse_sd(th))
    assert np.isnan(A.mcse_mean(bern["theta"][:3]))


def test_autocovariance_ar1():
    y = np.loadtxt(os.path.join(GOLDEN, "ar1.csv"))
    assert y.size == 1000
    ac = A.autocovariance(y)
    for lag, want in enumerate([2.69, 2.16, 1.73, 1.38, 1.10, 0.89]):
        assert abs(ac[lag] - want) < 0.01


def test_split_chains_odd_rows():
    foo = np.arange(1, 15, dtype=float).reshape(2, 7).T  # 7 x 2, column-major fill
    bar = A.split_chains(foo)
    assert bar.shape == (3, 4)
    np.testing.assert_array_equal(bar, [[1, 5, 8, 12], [2, 6, 9, 13], [3, 7, 10, 14]])


def test_min_bulk_ess_matches_per_parameter(bern):
    """The batched estimator equals split_rank_normalized_ess per parameter,
    including ties (repeated draws) and parameters that do not vary."""
    rng = np.random.Generator(np.random.PCG64(3))
    n, m, P = 401, 4, 24
    x = np.zeros((n, m, P))
    for t in range(1, n):
        x[t] = 0.7 * x[t - 1] + rng.standard_normal((m, P))
    x[5:9, 1, 3] = x[4, 1, 3]
    x[:, :, 7] = np.round(x[:, :, 7], 1)
    x[:, 2, 9] = 1.0
    per = np.array([A.split_rank_normalized_ess(x[:, :, p])[0] for p in range(P)])
    assert np.isnan(per[9])
    for block, threads in [(256, 1), (5, 3)]:
        got = A.min_bulk_ess(x, block=block, threads=threads)
        assert abs(got - np.nanmin(per)) <= 1e-9 * np.nanmin(per)
    # the reference's lp__ / theta columns, as a 2-parameter draw array
    d = np.stack([bern["lp"], bern["theta"]], axis=2)
    assert abs(A.min_bulk_ess(d) - 1407.5124) < 1e-4
    assert np.isnan(A.min_bulk_ess(x[:, :, [9]]))


def test_native_ess_matches_goldens(bern):
    """The C++ port used by the serving loop (stanx/analyze.hpp)."""
    import stan_amd as S
    assert abs(S.native_ess(bern["lp"], "ess") - 1335.41366787) < 1e-8
    assert abs(S.native_ess(bern["theta"], "ess") - 1377.5030282) < 1e-8
    assert abs(S.native_ess(bern["lp"]) - 1512.7684) < 1e-4
    assert abs(S.native_ess(bern["theta"]) - 1407.5124) < 1e-4
    assert np.isnan(S.native_ess(bern["divergent"]))
    rng = np.random.Generator(np.random.PCG64(9))
    x = np.zeros((301, 3))
    for t in range(1, 301):
        x[t] = 0.6 * x[t - 1] + rng.standard_normal(3)
    x[10:14, 2] = x[9, 2]
    assert abs(S.native_ess(x) - A.split_rank_normalized_ess(x)[0]) < 1e-9 * S.native_ess(x)
