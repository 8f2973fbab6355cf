"""Posterior draw diagnostics used to report min-ESS/s and MCSE.

Restates, with numpy, the reference's analysis routines (paths relative to
the stan-dev/stan 2.35.0 tree, src/stan/analyze/mcmc/):
  autocovariance           autocovariance.hpp:39-100 (FFT, biased 1/N estimator)
  ess                      ess.hpp:28-103 (Geyer initial monotone sequence)
  split_chains             split_chains.hpp:49-67
  rank_transform           rank_normalization.hpp:23-57 (Blom offsets 3/8)
  split_rank_normalized_ess split_rank_normalized_ess.hpp:28-55 (bulk, tail)
  mcse_mean / mcse_sd      mcse.hpp:23-62
  is_finite_and_varies     check_chains.hpp:20-38
Chains are columns: draws has shape (iterations, chains).
"""
import numpy as np
from scipy.special import ndtri


def _approx_const(v, ref):
    # Eigen isApproxToConstant: |v - ref| <= 1e-12 * min(|v|, |ref|) elementwise
    return np.all(np.abs(v - ref) <= 1e-12 * np.minimum(np.abs(v), abs(ref)))


def is_finite_and_varies(chains):
    chains = np.asarray(chains, dtype=np.float64)
    if not np.all(np.isfinite(chains)):
        return False
    first = chains[0, :]
    for i in range(chains.shape[1]):
        if _approx_const(chains[:, i], first[i]):
            return False
    if chains.shape[1] > 1 and _approx_const(first, first[0]):
        return False
    return True


def autocovariance(y):
    """Biased (1/N) autocovariance at every lag, via a zero-padded FFT."""
    y = np.asarray(y, dtype=np.float64)
    N = y.size
    M = 1
    while M < N:
        M <<= 1
    c = np.zeros(2 * M)
    c[:N] = y - y.mean()
    f = np.fft.rfft(c)
    ac = np.fft.irfft(f * np.conj(f), n=2 * M)[:N]
    ac = ac / ac[0]
    return ac * np.var(y)  # population variance (Boost accumulator)


def ess(chains):
    chains = np.asarray(chains, dtype=np.float64)
    n, m = chains.shape
    acov = np.column_stack([autocovariance(chains[:, i]) for i in range(m)])
    chain_mean = chains.mean(axis=0)
    chain_var = acov[0, :] * n / (n - 1)
    w = chain_var.mean()
    var_plus = w * (n - 1) / n
    if m > 1:
        var_plus += np.var(chain_mean, ddof=1)
    rho = np.zeros(n)
    rho_even = 1.0
    rho[0] = rho_even
    rho_odd = 1 - (w - acov[1, :].mean()) / var_plus
    rho[1] = rho_odd
    t = 1
    while t < n - 4 and (rho_even + rho_odd > 0) and not np.isnan(rho_even + rho_odd):
        rho_even = 1 - (w - acov[t + 1, :].mean()) / var_plus
        rho_odd = 1 - (w - acov[t + 2, :].mean()) / var_plus
        if rho_even + rho_odd >= 0:
            rho[t + 1] = rho_even
            rho[t + 2] = rho_odd
        if rho[t + 1] + rho[t + 2] > rho[t - 1] + rho[t]:
            rho[t + 1] = (rho[t - 1] + rho[t]) / 2
            rho[t + 2] = rho[t + 1]
        t += 2
    max_t = t
    if rho_even > 0:
        rho[max_t + 1] = rho_even
    total = m * n
    tau = -1 + 2 * rho[:max_t].sum() + rho[max_t + 1]
    tau = max(tau, 1 / np.log10(total))
    return total / tau


def split_chains(chains):
    chains = np.asarray(chains, dtype=np.float64)
    n = chains.shape[0]
    half = n // 2
    tail = (n + 1) // 2
    cols = []
    for i in range(chains.shape[1]):
        cols.append(chains[:half, i])
        cols.append(chains[tail:tail + half, i])
    return np.column_stack(cols)


def rank_transform(chains):
    chains = np.asarray(chains, dtype=np.float64)
    flat = chains.flatten(order="F")
    size = flat.size
    order = np.argsort(flat, kind="stable")
    sorted_v = flat[order]
    ranks = np.empty(size)
    i = 0
    while i < size:
        j = i + 1
        while j < size and sorted_v[j] == sorted_v[i]:
            j += 1
        avg = (i + 1 + j) / 2.0  # mean of ranks i+1 .. j
        ranks[order[i:j]] = avg
        i = j
    z = ndtri((ranks - 0.375) / (size + 0.25))
    return z.reshape(chains.shape, order="F")


def split_rank_normalized_ess(chains):
    """(bulk ESS, tail ESS)."""
    s = split_chains(chains)
    if not is_finite_and_varies(s) or s.shape[0] < 4:
        return float("nan"), float("nan")
    bulk = ess(rank_transform(s))
    flat = s.flatten(order="F")
    q05 = np.quantile(flat, 0.05)
    q95 = np.quantile(flat, 0.95)
    e05 = ess((s <= q05).astype(np.float64))
    e95 = ess((s >= q95).astype(np.float64))
    if np.isnan(e05):
        tail = e95
    elif np.isnan(e95):
        tail = e05
    else:
        tail = min(e05, e95)
    return bulk, tail


def mcse_mean(chains):
    chains = np.asarray(chains, dtype=np.float64)
    if chains.shape[0] < 4 or not is_finite_and_varies(chains):
        return float("nan")
    var = ((chains - chains.mean()) ** 2).sum() / (chains.size - 1)
    return float(np.sqrt(var / ess(chains)))


def mcse_sd(chains):
    chains = np.asarray(chains, dtype=np.float64)
    if chains.shape[0] < 4 or not is_finite_and_varies(chains):
        return float("nan")
    ctr = np.abs(chains - chains.mean())
    ess_mean = ess(split_chains(ctr))
    Evar = (ctr ** 2).mean()
    varvar = ((ctr ** 4).mean() - Evar ** 2) / ess_mean
    return float(np.sqrt(varvar / Evar / 4.0))


def min_bulk_ess(draws):
    """Minimum bulk ESS over parameters; draws has shape
    (iterations, chains, parameters)."""
    draws = np.asarray(draws, dtype=np.float64)
    vals = [split_rank_normalized_ess(draws[:, :, p])[0] for p in range(draws.shape[2])]
    vals = [v for v in vals if np.isfinite(v)]
    return float(min(vals)) if vals else float("nan")
