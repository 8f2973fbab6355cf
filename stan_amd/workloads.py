"""Synthetic data sets of the benchmark configurations (deterministic).

Each generator is seeded with numpy's PCG64 so the same arrays come out on
every machine running this image.  Sizes follow the configurations the
benchmarks are quoted on:

  logistic_small   N=10,000  K=50   X ~ N(0,1),    beta ~ N(0,1), alpha=0.5
  logistic_large   N=1e6     K=100  X ~ N(0,1/K),  beta ~ N(0,1), alpha=0.5
  normal_batched   N=200,000 K=256  X ~ N(0,1/K),  sigma=1, alpha=0.5
  hier_logistic    J=1000 groups, sizes round(LogNormal(ln 400, 1)), K=20
"""
import numpy as np

LOGIT, NORMAL, HIER_LOGIT = 0, 1, 2

# priors (alpha_sd, beta_sd, sigma_sd, tau_sd); alpha_sd <= 0 is flat
PRIOR_LOGISTIC = (5.0, 2.0, 2.0, 1.0)     # beta ~ N(0,2), alpha ~ N(0,5)
PRIOR_NORMAL = (0.0, 1.0, 2.0, 1.0)       # beta ~ N(0,1), sigma ~ half-N(0,2), alpha flat
PRIOR_HIER = (5.0, 2.0, 2.0, 1.0)         # tau ~ half-N(0,1), z ~ N(0,1)


def logistic(N, K, x_scale=1.0, alpha=0.5, seed=1234):
    rng = np.random.Generator(np.random.PCG64(seed))
    X = rng.standard_normal((N, K)) * x_scale
    beta = rng.standard_normal(K)
    eta = X @ beta + alpha
    p = 1.0 / (1.0 + np.exp(-eta))
    y = (rng.random(N) < p).astype(np.int32)
    return dict(family=LOGIT, X=X, y=y, beta=beta, alpha=alpha, prior=PRIOR_LOGISTIC)


def logistic_small(seed=1234):
    return logistic(10_000, 50, 1.0, 0.5, seed)


def logistic_large(N=1_000_000, K=100, seed=20250101):
    return logistic(N, K, 1.0 / np.sqrt(K), 0.5, seed)


def normal(N, K, sigma=1.0, alpha=0.5, seed=20250202):
    rng = np.random.Generator(np.random.PCG64(seed))
    X = rng.standard_normal((N, K)) / np.sqrt(K)
    beta = rng.standard_normal(K)
    y = X @ beta + alpha + sigma * rng.standard_normal(N)
    return dict(family=NORMAL, X=X, y=y, beta=beta, alpha=alpha, sigma=sigma,
                prior=PRIOR_NORMAL)


def normal_batched(seed=20250202):
    return normal(200_000, 256, 1.0, 0.5, seed)


def hier_logistic(J=1000, K=20, seed=20250303, mean_log_size=np.log(400.0), sd_log=1.0):
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = np.maximum(1, np.round(rng.lognormal(mean_log_size, sd_log, J))).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    N = int(offsets[-1])
    X = rng.standard_normal((N, K)) / np.sqrt(K)
    beta = rng.standard_normal(K)
    tau = np.abs(rng.standard_normal(K + 1))
    z = rng.standard_normal((J, K + 1))
    alpha = 0.5
    g = np.repeat(np.arange(J), sizes)
    eta = alpha + tau[0] * z[g, 0] + np.einsum("nk,nk->n", X, beta[None, :] + tau[None, 1:] * z[g, 1:])
    p = 1.0 / (1.0 + np.exp(-eta))
    y = (rng.random(N) < p).astype(np.int32)
    return dict(family=HIER_LOGIT, X=X, y=y, offsets=offsets, beta=beta, tau=tau, z=z,
                alpha=alpha, prior=PRIOR_HIER)


def random_thetas(D, count, seed=99):
    """Test points theta ~ N(0, 1) in R^D."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal((count, D))
