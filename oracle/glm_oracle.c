/*
 * glm_oracle.c -- test infrastructure (checker only): FP64 log density and
 * gradient of the three GLM models on the hot path.
 *
 * The reference evaluates these through Stan Math (absent from the tree:
 * lib/stan_math is an empty submodule, .gitmodules:1-3), reached from the
 * sampler via src/stan/model/gradient.hpp:21-35 with propto=true and
 * jacobian=true (src/stan/model/model_functional.hpp:21).  The formulas
 * are restated from the published definitions of
 *   bernoulli_logit_glm_lpmf(y | X, alpha, beta)
 *       = sum_n y_n eta_n - log1p(exp(eta_n)),   eta = X beta + alpha
 *       d/d eta_n = y_n - inv_logit(eta_n)
 *   normal_id_glm_lpdf(y | X, alpha, beta, sigma)
 *       = sum_n -0.5 ((y_n - mu_n)/sigma)^2 - log sigma [- 0.5 log 2 pi]
 * with normal priors (normal_lpdf) and the lower-bound transform
 * sigma = exp(u) whose log-Jacobian is +u (src/stan/io/deserializer.hpp:384-387).
 * The specification call sites are src/test/test-models/performance/logistic.stan:13-14,
 * src/test/test-models/good/services/normal_glm.stan:27 and, for the
 * hierarchical model, src/test/test-models/good/variational/hier_logistic.stan.
 *
 * Mode 0 restates the two-pass evaluation a CPU GLM performs (GEMV for
 * eta, elementwise link, GEMV X^T r) in plain double: it is the CPU
 * baseline.  Mode 1 uses error-free transformations (Knuth two-sum, FMA
 * two-product; Ogita, Rump and Oishi 2005 "Dot2"/"Sum2") so that its own
 * rounding is far below 1e-12 relative and summation order is not the
 * error source when the GPU kernel is compared against it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define HALF_LOG_2PI 0.91893853320467274178032973640562

/* ---- double-double accumulator ---- */
typedef struct { double s, c; } dd;

static inline void two_sum(double a, double b, double* s, double* e) {
  double x = a + b;
  double bv = x - a;
  double av = x - bv;
  *s = x;
  *e = (a - av) + (b - bv);
}
static inline void dd_add(dd* acc, double v) {
  double s, e;
  two_sum(acc->s, v, &s, &e);
  acc->s = s;
  acc->c += e;
}
static inline void dd_add_prod(dd* acc, double a, double b) {
  double p = a * b;
  double pe = fma(a, b, -p);
  dd_add(acc, p);
  acc->c += pe;
}
static inline double dd_val(const dd* acc) { return acc->s + acc->c; }

/* log(1 + exp(x)) without overflow */
static double log1p_exp(double x) {
  return x > 0 ? x + log1p(exp(-x)) : log1p(exp(x));
}

int64_t orc_glm_num_params(int family, int64_t K, int64_t J) {
  switch (family) {
    case ORC_LOGIT: return K + 1;
    case ORC_NORMAL: return K + 2;
    case ORC_HIER: return 2 * K + 2 + J * (K + 1);
  }
  return -1;
}

/* normal_lpdf(x | 0, s) with or without constants; flat when s <= 0 */
static void normal_prior(double x, double s, int propto, dd* lp, double* g) {
  if (!(s > 0)) return;
  double zz = x / s;
  dd_add(lp, -0.5 * zz * zz);
  if (!propto) dd_add(lp, -log(s) - HALF_LOG_2PI);
  *g += -x / (s * s);
}

/* Bernoulli-logit per-row term: returns lp_n, sets r_n = y - inv_logit(eta) */
static inline double logit_row(double eta, int y, double* r) {
  double s = y ? 1.0 : -1.0;
  double se = s * eta;
  *r = s / (1.0 + exp(se));
  return -log1p_exp(-se);
}

static int logit_lpgrad(int mode, const double* X, const int32_t* y,
                        int64_t N, int64_t K, const orc_prior* pr,
                        const double* th, int propto, double* lp,
                        double* grad) {
  const double alpha = th[0];
  const double* beta = th + 1;
  if (mode == 0) {
    double* r = (double*)malloc(sizeof(double) * (N > 0 ? N : 1));
    double lpsum = 0.0;
    for (int64_t n = 0; n < N; ++n) {
      const double* x = X + n * K;
      double eta = 0.0;
      for (int64_t k = 0; k < K; ++k) eta += x[k] * beta[k];
      eta += alpha;
      if (!isfinite(eta)) { free(r); return 1; }
      lpsum += logit_row(eta, y[n], &r[n]);
    }
    double ga = 0.0;
    for (int64_t k = 0; k < K; ++k) grad[1 + k] = 0.0;
    for (int64_t n = 0; n < N; ++n) {
      const double* x = X + n * K;
      for (int64_t k = 0; k < K; ++k) grad[1 + k] += x[k] * r[n];
      ga += r[n];
    }
    grad[0] = ga;
    free(r);
    dd acc = {lpsum, 0.0};
    normal_prior(alpha, pr->alpha_sd, propto, &acc, &grad[0]);
    for (int64_t k = 0; k < K; ++k)
      normal_prior(beta[k], pr->beta_sd, propto, &acc, &grad[1 + k]);
    *lp = dd_val(&acc);
  } else {
    dd* g = (dd*)calloc((size_t)K + 1, sizeof(dd));
    dd acc = {0.0, 0.0};
    for (int64_t n = 0; n < N; ++n) {
      const double* x = X + n * K;
      dd e = {alpha, 0.0};
      for (int64_t k = 0; k < K; ++k) dd_add_prod(&e, x[k], beta[k]);
      double eta = dd_val(&e);
      if (!isfinite(eta)) { free(g); return 1; }
      double r;
      dd_add(&acc, logit_row(eta, y[n], &r));
      for (int64_t k = 0; k < K; ++k) dd_add_prod(&g[1 + k], x[k], r);
      dd_add(&g[0], r);
    }
    for (int64_t k = 0; k <= K; ++k) grad[k] = dd_val(&g[k]);
    free(g);
    normal_prior(alpha, pr->alpha_sd, propto, &acc, &grad[0]);
    for (int64_t k = 0; k < K; ++k)
      normal_prior(beta[k], pr->beta_sd, propto, &acc, &grad[1 + k]);
    *lp = dd_val(&acc);
  }
  return isfinite(*lp) ? 0 : 1;
}

static int normal_lpgrad(int mode, const double* X, const double* y,
                         int64_t N, int64_t K, const orc_prior* pr,
                         const double* th, int propto, int jacobian,
                         double* lp, double* grad) {
  const double alpha = th[0];
  const double* beta = th + 1;
  const double u = th[K + 1];
  const double sigma = exp(u);
  if (!(sigma > 0) || !isfinite(sigma)) return 1;
  const double inv_s2 = 1.0 / (sigma * sigma);
  dd ss = {0.0, 0.0}, acc = {0.0, 0.0};
  if (mode == 0) {
    double* r = (double*)malloc(sizeof(double) * (N > 0 ? N : 1));
    double ssum = 0.0;
    for (int64_t n = 0; n < N; ++n) {
      const double* x = X + n * K;
      double mu = 0.0;
      for (int64_t k = 0; k < K; ++k) mu += x[k] * beta[k];
      mu += alpha;
      double d = y[n] - mu;
      if (!isfinite(d)) { free(r); return 1; }
      r[n] = d;
      ssum += d * d;
    }
    double ga = 0.0;
    for (int64_t k = 0; k < K; ++k) grad[1 + k] = 0.0;
    for (int64_t n = 0; n < N; ++n) {
      const double* x = X + n * K;
      for (int64_t k = 0; k < K; ++k) grad[1 + k] += x[k] * r[n];
      ga += r[n];
    }
    grad[0] = ga * inv_s2;
    for (int64_t k = 0; k < K; ++k) grad[1 + k] *= inv_s2;
    free(r);
    ss.s = ssum;
  } else {
    dd* g = (dd*)calloc((size_t)K + 1, sizeof(dd));
    for (int64_t n = 0; n < N; ++n) {
      const double* x = X + n * K;
      dd e = {y[n], 0.0};
      dd_add(&e, -alpha);
      for (int64_t k = 0; k < K; ++k) dd_add_prod(&e, -x[k], beta[k]);
      double d = dd_val(&e);
      if (!isfinite(d)) { free(g); return 1; }
      dd_add_prod(&ss, d, d);
      for (int64_t k = 0; k < K; ++k) dd_add_prod(&g[1 + k], x[k], d);
      dd_add(&g[0], d);
    }
    for (int64_t k = 0; k <= K; ++k) grad[k] = dd_val(&g[k]) * inv_s2;
    free(g);
  }
  double SS = dd_val(&ss);
  dd_add(&acc, -0.5 * SS * inv_s2);
  dd_add(&acc, -(double)N * u);
  if (!propto) dd_add(&acc, -(double)N * HALF_LOG_2PI);
  grad[K + 1] = -(double)N + SS * inv_s2;
  normal_prior(alpha, pr->alpha_sd, propto, &acc, &grad[0]);
  for (int64_t k = 0; k < K; ++k)
    normal_prior(beta[k], pr->beta_sd, propto, &acc, &grad[1 + k]);
  /* sigma ~ normal(sigma_mu, sigma_sd) on sigma = exp(u):
     d/du = -(sigma - mu) sigma / sd^2 */
  if (pr->sigma_sd > 0) {
    double zz = (sigma - pr->sigma_mu) / pr->sigma_sd;
    dd_add(&acc, -0.5 * zz * zz);
    if (!propto) dd_add(&acc, -log(pr->sigma_sd) - HALF_LOG_2PI);
    grad[K + 1] += -zz * sigma / pr->sigma_sd;
  }
  if (jacobian) {
    dd_add(&acc, u);
    grad[K + 1] += 1.0;
  }
  *lp = dd_val(&acc);
  return isfinite(*lp) ? 0 : 1;
}

static int hier_lpgrad(int mode, const double* X, const int32_t* y,
                       int64_t N, int64_t K, const int64_t* off, int64_t J,
                       const orc_prior* pr, const double* th, int propto,
                       int jacobian, double* lp, double* grad) {
  const int64_t K1 = K + 1;
  const double alpha = th[0];
  const double* beta = th + 1;
  const double* v = th + 1 + K;
  const double* z = th + 2 + 2 * K;
  const int64_t D = orc_glm_num_params(ORC_HIER, K, J);
  double* tau = (double*)malloc(sizeof(double) * K1);
  double* w = (double*)malloc(sizeof(double) * K1);
  dd* S = (dd*)malloc(sizeof(dd) * K1);
  for (int64_t k = 0; k < K1; ++k) tau[k] = exp(v[k]);
  for (int64_t d = 0; d < D; ++d) grad[d] = 0.0;
  dd acc = {0.0, 0.0};
  double lpsum = 0.0;
  int status = 0;
  (void)N;
  for (int64_t j = 0; j < J && !status; ++j) {
    const double* zj = z + j * K1;
    double b = alpha + tau[0] * zj[0];
    for (int64_t k = 0; k < K; ++k) w[k] = beta[k] + tau[k + 1] * zj[k + 1];
    for (int64_t k = 0; k < K1; ++k) { S[k].s = 0.0; S[k].c = 0.0; }
    for (int64_t n = off[j]; n < off[j + 1]; ++n) {
      const double* x = X + n * K;
      double eta;
      if (mode == 0) {
        eta = 0.0;
        for (int64_t k = 0; k < K; ++k) eta += x[k] * w[k];
        eta += b;
      } else {
        dd e = {b, 0.0};
        for (int64_t k = 0; k < K; ++k) dd_add_prod(&e, x[k], w[k]);
        eta = dd_val(&e);
      }
      if (!isfinite(eta)) { status = 1; break; }
      double r;
      double t = logit_row(eta, y[n], &r);
      if (mode == 0) {
        lpsum += t;
        S[0].s += r;
        for (int64_t k = 0; k < K; ++k) S[k + 1].s += x[k] * r;
      } else {
        dd_add(&acc, t);
        dd_add(&S[0], r);
        for (int64_t k = 0; k < K; ++k) dd_add_prod(&S[k + 1], x[k], r);
      }
    }
    if (status) break;
    double* gz = grad + 2 + 2 * K + j * K1;
    for (int64_t k = 0; k < K1; ++k) {
      double s = dd_val(&S[k]);
      if (k == 0) grad[0] += s; else grad[k] += s; /* alpha / beta_k */
      gz[k] = tau[k] * s;
      grad[1 + K + k] += tau[k] * zj[k] * s; /* d/dv_k before prior */
    }
  }
  if (!status) {
    if (mode == 0) dd_add(&acc, lpsum);
    normal_prior(alpha, pr->alpha_sd, propto, &acc, &grad[0]);
    for (int64_t k = 0; k < K; ++k)
      normal_prior(beta[k], pr->beta_sd, propto, &acc, &grad[1 + k]);
    for (int64_t k = 0; k < K1; ++k) {
      if (pr->tau_sd > 0) {
        double zz = tau[k] / pr->tau_sd;
        dd_add(&acc, -0.5 * zz * zz);
        if (!propto) dd_add(&acc, -log(pr->tau_sd) - HALF_LOG_2PI);
        grad[1 + K + k] += -zz * zz;
      }
      if (jacobian) {
        dd_add(&acc, v[k]);
        grad[1 + K + k] += 1.0;
      }
    }
    for (int64_t i = 0; i < J * K1; ++i) {
      double zi = z[i];
      dd_add(&acc, -0.5 * zi * zi);
      if (!propto) dd_add(&acc, -HALF_LOG_2PI);
      grad[2 + 2 * K + i] += -zi;
    }
    *lp = dd_val(&acc);
    if (!isfinite(*lp)) status = 1;
  }
  free(tau); free(w); free(S);
  return status;
}

int orc_glm_lpgrad(int family, int mode, const double* X, const void* y,
                   int64_t N, int64_t K, const int64_t* offsets, int64_t J,
                   const orc_prior* prior, const double* theta, int propto,
                   int jacobian, double* lp, double* grad) {
  switch (family) {
    case ORC_LOGIT:
      return logit_lpgrad(mode, X, (const int32_t*)y, N, K, prior, theta,
                          propto, lp, grad);
    case ORC_NORMAL:
      return normal_lpgrad(mode, X, (const double*)y, N, K, prior, theta,
                           propto, jacobian, lp, grad);
    case ORC_HIER:
      return hier_lpgrad(mode, X, (const int32_t*)y, N, K, offsets, J, prior,
                         theta, propto, jacobian, lp, grad);
  }
  return 2;
}

/* ---- model callbacks for the oracle sampler ---- */
typedef struct orc_glm_ctx {
  int family, mode;
  const double* X;
  const void* y;
  int64_t N, K, J;
  const int64_t* offsets;
  orc_prior prior;
} orc_glm_ctx;

int orc_glm_model_cb(void* vctx, const double* q, int D, double* lp,
                     double* grad) {
  const orc_glm_ctx* c = (const orc_glm_ctx*)vctx;
  (void)D;
  return orc_glm_lpgrad(c->family, c->mode, c->X, c->y, c->N, c->K,
                        c->offsets, c->J, &c->prior, q, 1, 1, lp, grad);
}

/* gauss3D / command test models: V = |q|^2 / 2 (src/test/test-models/good/
 * mcmc/hmc/common/gauss3D.stan, .../integrators/command.stan at y=0) */
int orc_gauss_model_cb(void* ctx, const double* q, int D, double* lp,
                       double* grad) {
  (void)ctx;
  double s = 0.0;
  for (int i = 0; i < D; ++i) {
    s += q[i] * q[i];
    grad[i] = -q[i];
  }
  *lp = -0.5 * s;
  return 0;
}

/* funnel (src/test/test-models/good/mcmc/hmc/hamiltonians/funnel.stan):
 * v ~ normal(0, 3), x_i ~ normal(0, exp(v)), constants dropped */
int orc_funnel_model_cb(void* ctx, const double* q, int D, double* lp,
                        double* grad) {
  (void)ctx;
  double v = q[0];
  double lpv = -0.5 * v * v / 9.0;
  double gv = -v / 9.0;
  double e2 = exp(-2.0 * v);
  for (int i = 1; i < D; ++i) {
    lpv += -0.5 * q[i] * q[i] * e2 - v;
    gv += q[i] * q[i] * e2 - 1.0;
    grad[i] = -q[i] * e2;
  }
  grad[0] = gv;
  *lp = lpv;
  return 0;
}

/* Correlated Gaussian test model: x ~ N(0, S) with S_ij = s_i s_j rho^|i-j|
 * (AR(1) correlation, rho = 0.9, scales s_i = 1 + i), whose precision is
 * tridiagonal: with z_i = x_i / s_i, -2 log p = (sum_i z_i^2 (1 + rho^2 c_i)
 * - 2 rho sum_i z_i z_{i+1}) / (1 - rho^2), c_i = 1 for interior i, 0 at the
 * two ends.  A dense metric equal to S makes the problem isotropic. */
int orc_corr_gauss_model_cb(void* ctx, const double* q, int D, double* lp,
                            double* grad) {
  (void)ctx;
  const double rho = 0.9, w = 1.0 / (1.0 - rho * rho);
  double s = 0.0;
  for (int i = 0; i < D; ++i) {
    const double si = 1.0 + i, zi = q[i] / si;
    const double di = (i == 0 || i == D - 1) ? 1.0 : 1.0 + rho * rho;
    double gz = di * zi;
    s += di * zi * zi;
    if (i + 1 < D) {
      const double zn = q[i + 1] / (2.0 + i);
      s -= 2.0 * rho * zi * zn;
      gz -= rho * zn;
    }
    if (i > 0) gz -= rho * q[i - 1] / (double)i;
    grad[i] = -w * gz / si;
  }
  *lp = -0.5 * w * s;
  return 0;
}
