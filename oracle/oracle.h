/*
 * oracle.h -- CPU restatement of the reference's hot path, used ONLY as a
 * checker (by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg).  Nothing in the product library links or calls this.
 *
 * What is restated, and from where (paths relative to the stan-dev/stan
 * 2.35.0 tree):
 *   - NUTS diag_e transition and recursive build_tree:
 *       src/stan/mcmc/hmc/nuts/base_nuts.hpp:78-204, 247-352
 *   - explicit leapfrog: src/stan/mcmc/hmc/integrators/base_leapfrog.hpp:17-22,
 *       expl_leapfrog.hpp:16-32
 *   - diag_e metric: src/stan/mcmc/hmc/hamiltonians/diag_e_metric.hpp:20-50
 *   - potential/gradient with domain-error rejection:
 *       src/stan/mcmc/hmc/hamiltonians/base_hamiltonian.hpp:48-70
 *   - step-size heuristic: src/stan/mcmc/hmc/base_hmc.hpp:93-161, 195-200
 *   - dual averaging: src/stan/mcmc/stepsize_adaptation.hpp:49-73
 *   - windowed variance adaptation: src/stan/mcmc/windowed_adaptation.hpp:23-110,
 *       src/stan/mcmc/var_adaptation.hpp:17-46 (Welford estimator from Stan
 *       Math, absent; restated from Welford 1962)
 *   - adaptive transition: src/stan/mcmc/hmc/nuts/adapt_diag_e_nuts.hpp:25-48
 *   - dense_e metric: src/stan/mcmc/hmc/hamiltonians/dense_e_metric.hpp:21-54
 *       (Cholesky factor and triangular solve restated from their
 *       definitions; Eigen's LLT is absent), windowed covariance adaptation
 *       src/stan/mcmc/covar_adaptation.hpp:17-47 (Welford covariance
 *       estimator from Stan Math, absent), adaptive transition
 *       src/stan/mcmc/hmc/nuts/adapt_dense_e_nuts.hpp:25-48, metric checks
 *       src/stan/services/util/validate_dense_inv_metric.hpp:18-28
 *   - service driver: src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:57-112,
 *       src/stan/services/sample/hmc_nuts_dense_e_adapt.hpp:57-112,
 *       src/stan/services/sample/hmc_nuts_dense_e.hpp:48-93,
 *       src/stan/services/util/run_sampler.hpp:39-85,
 *       src/stan/services/util/run_adaptive_sampler.hpp:46-103,
 *       src/stan/services/util/generate_transitions.hpp:42-77,
 *       src/stan/services/util/initialize.hpp:68-227
 *   - GLM likelihoods (Stan Math bernoulli_logit_glm_lpmf / normal_id_glm_lpdf,
 *       absent from the tree; restated from their published mathematics):
 *       specification call sites src/test/test-models/performance/logistic.stan:13-14,
 *       src/test/test-models/good/services/normal_glm.stan:27
 *
 * The RNG is Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), not Boost's
 * mixmax: trajectories therefore match the reference statistically, not
 * bitwise.  The same counter-based stream is restated independently in the
 * product so that product and oracle runs can be compared draw for draw.
 */
#ifndef STANX_ORACLE_H
#define STANX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- RNG ---------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);

typedef struct orc_rng {
  uint32_t key[2];
  uint64_t counter;
} orc_rng;

void orc_rng_init(orc_rng* r, uint32_t seed, uint32_t chain);
double orc_rng_uniform(orc_rng* r); /* [0,1) */
double orc_rng_normal(orc_rng* r);  /* N(0,1), Box-Muller (cos branch) */

/* ---------------- GLM log density + gradient ---------------- */
/* Families; theta layouts (unconstrained):
 *   LOGIT : [alpha, beta_1..beta_K]                          D = K+1
 *   NORMAL: [alpha, beta_1..beta_K, u]  sigma = exp(u)       D = K+2
 *   HIER  : [alpha, beta_1..K, v_0..v_K, z_{1,0..K}..z_{J,0..K}]
 *           tau_k = exp(v_k)                                 D = 2K+2+J(K+1)
 */
enum { ORC_LOGIT = 0, ORC_NORMAL = 1, ORC_HIER = 2 };

typedef struct orc_prior {
  double alpha_sd; /* alpha ~ normal(0, alpha_sd); <= 0 means flat */
  double beta_sd;  /* beta  ~ normal(0, beta_sd) */
  double sigma_sd; /* sigma ~ half-normal(0, sigma_sd)   (NORMAL) */
  double tau_sd;   /* tau   ~ half-normal(0, tau_sd)     (HIER)   */
  double sigma_mu; /* sigma ~ normal(sigma_mu, sigma_sd) on sigma > 0 */
} orc_prior;

/* mode 0: Stan-equivalent two-pass double evaluation (eta = X beta + alpha,
 *         then X^T r), plain double accumulation -- the CPU baseline.
 * mode 1: compensated evaluation (double-double dot products and sums) --
 *         the parity reference.
 * X is row-major N x K doubles.  y is int32 {0,1} for LOGIT/HIER and double
 * for NORMAL.  offsets (J+1 entries, CSR) only for HIER.
 * Returns 0, or 1 when the log density is not finite (domain error).   */
int orc_glm_lpgrad(int family, int mode, const double* X, const void* y,
                   int64_t N, int64_t K, const int64_t* offsets, int64_t J,
                   const orc_prior* prior, const double* theta, int propto,
                   int jacobian, double* lp, double* grad);

int64_t orc_glm_num_params(int family, int64_t K, int64_t J);

/* ---------------- NUTS ---------------- */
/* Model callback: value and gradient of log p(q) (propto, jacobian).
 * Returns 0 on success, 1 for a domain error (proposal rejected), any other
 * value for a fatal error.                                              */
typedef int (*orc_lpgrad_fn)(void* ctx, const double* q, int D, double* lp,
                             double* grad);

typedef struct orc_nuts_args {
  uint32_t seed;
  uint32_t chain;
  double init_radius;
  int num_warmup, num_samples, num_thin, save_warmup;
  double stepsize, stepsize_jitter;
  int max_depth;
  double delta, gamma, kappa, t0;
  unsigned init_buffer, term_buffer, window;
  const double* init;          /* optional D-vector of unconstrained inits */
  const double* init_inv_metric; /* optional: D-vector (diag), or D x D
                                    row-major (dense) */
} orc_nuts_args;

/* Draw table: one row per saved iteration, columns
 *   lp__, accept_stat__, stepsize__, treedepth__, n_leapfrog__,
 *   divergent__, energy__, q_1..q_D (unconstrained)
 * out_draws must hold (rows x (7 + D)) doubles.  Rows = number of saved
 * iterations.  Returns 0 (OK), 70 (SOFTWARE) or 78 (CONFIG) like the
 * reference service.  out_stepsize / out_inv_metric receive the adapted
 * values; n_grad counts every gradient evaluation.                    */
int orc_hmc_nuts_diag_e_adapt(orc_lpgrad_fn f, void* ctx, int D,
                              const orc_nuts_args* a, double* out_draws,
                              int64_t* out_rows, double* out_stepsize,
                              double* out_inv_metric, int64_t* n_grad);
/* hmc_nuts_diag_e: the same transitions without adaptation (no step-size
 * heuristic; stepsize and metric fixed). */
int orc_hmc_nuts_diag_e(orc_lpgrad_fn f, void* ctx, int D, const orc_nuts_args* a,
                        double* out_draws, int64_t* out_rows, int64_t* n_grad);
/* hmc_nuts_dense_e_adapt / hmc_nuts_dense_e: dense metric (D x D row-major
 * init_inv_metric, identity if NULL; out_inv_metric receives D x D). */
int orc_hmc_nuts_dense_e_adapt(orc_lpgrad_fn f, void* ctx, int D,
                               const orc_nuts_args* a, double* out_draws,
                               int64_t* out_rows, double* out_stepsize,
                               double* out_inv_metric, int64_t* n_grad);
int orc_hmc_nuts_dense_e(orc_lpgrad_fn f, void* ctx, int D, const orc_nuts_args* a,
                         double* out_draws, int64_t* out_rows, int64_t* n_grad);
/* Cholesky factor A = L L^T (row-major, lower L); 0, or -1 if A is not
 * positive definite. */
int orc_cholesky(const double* A, int D, double* L);

/* Single pieces exposed for golden tests. */
typedef struct orc_point {
  int D;
  double *q, *p, *g, *minv; /* minv: D (diag) or D x D row-major (dense) */
  double V;
  int dense;
} orc_point;

/* leapfrog on a point with a model callback; returns 0 or fatal code */
int orc_leapfrog(orc_lpgrad_fn f, void* ctx, orc_point* z, double eps);
double orc_hamiltonian(const orc_point* z);

/* build_tree with the exact argument semantics of base_nuts::build_tree.
 * All vectors are D long.  The RNG and eps are supplied explicitly.      */
int orc_build_tree(orc_lpgrad_fn f, void* ctx, orc_rng* rng, double eps,
                   double max_deltaH, int depth, orc_point* z,
                   orc_point* z_propose, double* p_sharp_beg,
                   double* p_sharp_end, double* rho, double* p_beg,
                   double* p_end, double H0, double sign, int* n_leapfrog,
                   double* log_sum_weight, double* sum_metro_prob,
                   int* divergent);

/* Dual averaging step (stepsize_adaptation::learn_stepsize). state =
 * {counter, s_bar, x_bar, mu, delta, gamma, kappa, t0}.                */
void orc_learn_stepsize(double* state, double* epsilon, double adapt_stat);
/* var_adaptation_test.cpp:5-27: the metric after n_learn zero draws */
int orc_var_window_zero(int D, int n_learn, unsigned num_warmup, unsigned init_buffer,
                        unsigned term_buffer, unsigned base_window, double* var);

#ifdef __cplusplus
}
#endif
#endif
