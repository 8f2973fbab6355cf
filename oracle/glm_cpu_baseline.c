/*
 * glm_cpu_baseline.c -- test infrastructure: the CPU leg the benchmark
 * times beside the GPU (never a checker of results and never part of the
 * product).
 *
 * The reference evaluates bernoulli_logit_glm_lpmf on the CPU through Stan
 * Math (absent from the tree) in two passes over X: eta = X beta + alpha,
 * the elementwise link, then the adjoint X^T r (the GEMV pair of a CPU
 * GLM).  This restates those two passes, with the same arithmetic as the
 * oracle's mode 0 (glm_oracle.c), parallel over rows: each thread takes a
 * contiguous block of rows for both passes and keeps its own gradient
 * partial; the partials are added in thread order at the end.  Built with
 * -O3 for a widely available x86-64 level and OpenMP (oracle/Makefile).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double log1p_exp(double x) { return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x)); }

/* lp (likelihood + normal priors, propto) and gradient at th = (alpha,
 * beta) with `threads` OpenMP threads; returns 1 on a non-finite eta. */
int cpu_logit_lpgrad(const double* X, const int32_t* y, int64_t N, int64_t K, double alpha_sd,
                     double beta_sd, const double* th, int threads, double* lp, double* grad) {
  const double alpha = th[0];
  const double* beta = th + 1;
  if (threads < 1) threads = 1;
  double* part = (double*)calloc((size_t)threads * (size_t)(K + 2), sizeof(double));
  int bad = 0;
#pragma omp parallel num_threads(threads) reduction(| : bad)
  {
    const int t = omp_get_thread_num();
    const int nt = omp_get_num_threads();
    const int64_t n0 = N * t / nt, n1 = N * (t + 1) / nt;
    double* g = part + (size_t)t * (size_t)(K + 2); /* [0]: lp, [1]: alpha, [2..]: beta */
    double lps = 0.0, ga = 0.0;
    const int64_t B = 256; /* rows per block: eta, then the adjoint of the block */
    double r[256];
    for (int64_t b0 = n0; b0 < n1; b0 += B) {
      const int64_t b1 = b0 + B < n1 ? b0 + B : n1;
      for (int64_t n = b0; n < b1; ++n) {
        const double* x = X + n * K;
        double eta = 0.0;
        for (int64_t k = 0; k < K; ++k) eta += x[k] * beta[k];
        eta += alpha;
        if (!isfinite(eta)) bad = 1;
        const double s = y[n] ? 1.0 : -1.0;
        const double se = s * eta;
        r[n - b0] = s / (1.0 + exp(se));
        lps += -log1p_exp(-se);
      }
      for (int64_t n = b0; n < b1; ++n) {
        const double* x = X + n * K;
        const double rn = r[n - b0];
        for (int64_t k = 0; k < K; ++k) g[2 + k] += x[k] * rn;
        ga += rn;
      }
    }
    g[0] = lps;
    g[1] = ga;
  }
  double l = 0.0;
  for (int64_t k = 0; k < K + 1; ++k) grad[k] = 0.0;
  for (int t = 0; t < threads; ++t) {
    const double* g = part + (size_t)t * (size_t)(K + 2);
    l += g[0];
    grad[0] += g[1];
    for (int64_t k = 0; k < K; ++k) grad[1 + k] += g[2 + k];
  }
  free(part);
  if (alpha_sd > 0) {
    l += -0.5 * (alpha / alpha_sd) * (alpha / alpha_sd);
    grad[0] += -alpha / (alpha_sd * alpha_sd);
  }
  if (beta_sd > 0)
    for (int64_t k = 0; k < K; ++k) {
      l += -0.5 * (beta[k] / beta_sd) * (beta[k] / beta_sd);
      grad[1 + k] += -beta[k] / (beta_sd * beta_sd);
    }
  *lp = l;
  return bad;
}
