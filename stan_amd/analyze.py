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
    with np.errstate(invalid="ignore", divide="ignore"):
        ac = ac / ac[0]
    return ac * np.var(y)  # population variance (Boost accumulator)


def _ess_core(acov, chain_mean, n, m):
    """Geyer initial monotone sequence estimator from the per-chain
    autocovariances acov (n x m) and chain means (ess.hpp:28-103)."""
    chain_var = acov[0, :] * n / (n - 1)
    w = chain_var.mean()
    var_plus = w * (n - 1) / n
    if m > 1:
        var_plus += np.var(chain_mean, ddof=1)
    return _ess_geyer(acov.mean(axis=1), w, var_plus, n, m)


def _ess_geyer(acov_mean, w, var_plus, n, m):
    """The initial monotone sequence part of ess() from the chain-averaged
    autocovariance, the mean within-chain variance w and var_plus."""
    rho = np.zeros(n)
    rho_even = 1.0
    rho[0] = rho_even
    rho_odd = 1 - (w - acov_mean[1]) / var_plus
    rho[1] = rho_odd
    t = 1
    while t < n - 4 and (rho_even + rho_odd > 0) and not np.isnan(rho_even + rho_odd):
        rho_even = 1 - (w - acov_mean[t + 1]) / var_plus
        rho_odd = 1 - (w - acov_mean[t + 2]) / var_plus
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


def _autocovariance_cols(y):
    """autocovariance() of every column of y (n x ...) at once."""
    n = y.shape[0]
    M = 1
    while M < n:
        M <<= 1
    c = np.zeros((2 * M,) + y.shape[1:])
    c[:n] = y - y.mean(axis=0)
    f = np.fft.rfft(c, axis=0)
    ac = np.fft.irfft(f * np.conj(f), n=2 * M, axis=0)[:n]
    with np.errstate(invalid="ignore", divide="ignore"):
        ac = ac / ac[0]
    return ac * np.var(y, axis=0)


def ess(chains):
    chains = np.asarray(chains, dtype=np.float64)
    n, m = chains.shape
    acov = np.column_stack([autocovariance(chains[:, i]) for i in range(m)])
    return _ess_core(acov, chains.mean(axis=0), n, m)


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


def _bulk_ess_block(x):
    """Bulk ESS of every parameter of x (params x 2m split chains x half
    iterations); NaN where the split chains are not finite or do not vary."""
    nb, m2, half = x.shape
    size = m2 * half
    first = x[:, :, 0]
    ok = np.isfinite(x).all(axis=(1, 2))
    with np.errstate(invalid="ignore"):
        # a chain approximately constant (Eigen isApproxToConstant); the
        # elementwise test only runs on chains whose spread is tiny
        spread = x.max(axis=2) - x.min(axis=2)
        cand = spread <= 1e-11 * np.abs(x).max(axis=2)
        for p, c in zip(*np.nonzero(cand)):
            row = x[p, c]
            if np.all(np.abs(row - row[0]) <= 1e-12 * np.minimum(np.abs(row), abs(row[0]))):
                ok[p] = False
        if m2 > 1:
            same0 = np.abs(first - first[:, 0:1]) <= 1e-12 * np.minimum(np.abs(first),
                                                                        np.abs(first[:, 0:1]))
            ok &= ~same0.all(axis=1)
    out = np.full(nb, np.nan)
    sel = np.nonzero(ok)[0]
    if sel.size == 0:
        return out
    # rank transform of each parameter's column-major (half x 2m) matrix:
    # average ranks over ties, Blom normal scores
    flat = x[sel].reshape(sel.size, size)
    order = np.argsort(flat, axis=1, kind="stable")
    v = np.take_along_axis(flat, order, axis=1)
    avg = np.broadcast_to(np.arange(1.0, size + 1.0), v.shape).copy()
    tied = v[:, 1:] == v[:, :-1]
    for r in np.nonzero(tied.any(axis=1))[0]:
        # mean of the ranks of each run of equal values
        ar = np.arange(size)
        new = np.ones(size, bool)
        new[1:] = ~tied[r]
        last = np.ones(size, bool)
        last[:-1] = new[1:]
        start = np.maximum.accumulate(np.where(new, ar, 0))
        end = np.minimum.accumulate(np.where(last, ar, size)[::-1])[::-1]
        avg[r] = (start + end + 2) / 2.0
    ranks = np.empty(v.shape)
    np.put_along_axis(ranks, order, avg, axis=1)
    z = ndtri((ranks - 0.375) / (size + 0.25)).reshape(sel.size, m2, half)
    M = 1
    while M < half:
        M <<= 1
    c = np.zeros((sel.size, m2, 2 * M))
    c[:, :, :half] = z - z.mean(axis=2, keepdims=True)
    f = np.fft.rfft(c, axis=2)
    ac = np.fft.irfft(f * np.conj(f), n=2 * M, axis=2)[:, :, :half]
    with np.errstate(invalid="ignore", divide="ignore"):
        ac = ac / ac[:, :, 0:1] * z.var(axis=2, keepdims=True)
    w = (ac[:, :, 0] * half / (half - 1)).mean(axis=1)
    var_plus = w * (half - 1) / half
    if m2 > 1:
        var_plus = var_plus + z.mean(axis=2).var(axis=1, ddof=1)
    acov_mean = ac.mean(axis=1)
    for i, j in enumerate(sel):
        out[j] = _ess_geyer(acov_mean[i], w[i], var_plus[i], half, m2)
    return out


def min_bulk_ess(draws, block=256, threads=6):
    """Minimum bulk ESS over parameters; draws has shape
    (iterations, chains, parameters).  Same estimator as
    split_rank_normalized_ess(draws[:, :, p])[0] per parameter (split
    chains, average ranks over ties, Blom normal scores, FFT
    autocovariances, Geyer's initial monotone sequence), vectorised over
    blocks of parameters that a thread pool works through, so that tens of
    thousands of parameters take seconds.  Parameters whose split chains
    are not finite or do not vary are skipped (split_rank_normalized_ess
    returns NaN for them)."""
    from concurrent.futures import ThreadPoolExecutor
    draws = np.asarray(draws, dtype=np.float64)
    n, m, P = draws.shape
    half, tail = n // 2, (n + 1) // 2
    if half < 4:
        return float("nan")

    def split(b0):
        # params x (chain 0 first half, chain 0 second half, ...) x iterations
        blk = draws[:, :, b0:b0 + block]
        x = np.empty((blk.shape[2], 2 * m, half))
        x[:, 0::2, :] = np.transpose(blk[:half], (2, 1, 0))
        x[:, 1::2, :] = np.transpose(blk[tail:tail + half], (2, 1, 0))
        return _bulk_ess_block(x)

    starts = range(0, P, block)
    if threads > 1 and P > block:
        with ThreadPoolExecutor(threads) as ex:
            vals = np.concatenate(list(ex.map(split, starts)))
    else:
        vals = np.concatenate([split(b) for b in starts])
    vals = vals[np.isfinite(vals)]
    return float(vals.min()) if vals.size else float("nan")
