/*
 * nuts_oracle.c -- test infrastructure (checker only): recursive multinomial
 * NUTS with a diagonal or dense Euclidean metric, its adaptation and the
 * service driver, restated line by line from the reference (files listed
 * in oracle.h).  This is the "reference sampler": posterior moments of the
 * product's samplers are compared with runs of this code, and its speed on
 * host cores is the CPU baseline.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define NEG_INF (-INFINITY)

/* Stan Math log1p_exp / log_sum_exp(a, b) (absent from the tree; restated
 * from their definitions: log(1+e^a) and log(e^a + e^b) with the -inf
 * short cut).                                                         */
static double log1p_exp(double a) {
  return a > 0.0 ? a + log1p(exp(-a)) : log1p(exp(a));
}
static double log_sum_exp(double a, double b) {
  if (a == NEG_INF) return b;
  if (a == INFINITY && b == INFINITY) return INFINITY;
  if (a > b) return a + log1p_exp(b - a);
  return b + log1p_exp(a - b);
}

static double dot(const double* a, const double* b, int D) {
  double s = 0.0;
  for (int i = 0; i < D; ++i) s += a[i] * b[i];
  return s;
}

/* ---------------- point helpers ---------------- */
static void pt_alloc(orc_point* z, int D, double* minv, int dense) {
  z->D = D;
  z->dense = dense;
  z->q = (double*)calloc((size_t)D + 1, sizeof(double));
  z->p = (double*)calloc((size_t)D + 1, sizeof(double));
  z->g = (double*)calloc((size_t)D + 1, sizeof(double));
  z->minv = minv;
  z->V = 0.0;
}
static void pt_free(orc_point* z) {
  free(z->q); free(z->p); free(z->g);
}
/* ps_point::operator= : copies q, p, g, V (not the metric) */
static void pt_copy(orc_point* dst, const orc_point* src) {
  size_t n = sizeof(double) * (size_t)src->D;
  memcpy(dst->q, src->q, n);
  memcpy(dst->p, src->p, n);
  memcpy(dst->g, src->g, n);
  dst->V = src->V;
}
static double* vec(int D) { return (double*)calloc((size_t)D + 1, sizeof(double)); }

/* p_sharp = M^-1 p: diag_e_metric::dtau_dp (diag_e_metric.hpp:36-38)
 * elementwise; dense_e_metric::dtau_dp (dense_e_metric.hpp:37) as a
 * matrix-vector product over the row-major matrix, j ascending.        */
static void dtau_dp(const orc_point* z, double* out) {
  if (z->dense) {
    for (int i = 0; i < z->D; ++i) {
      const double* row = z->minv + (size_t)i * (size_t)z->D;
      double s = 0.0;
      for (int j = 0; j < z->D; ++j) s += row[j] * z->p[j];
      out[i] = s;
    }
    return;
  }
  for (int i = 0; i < z->D; ++i) out[i] = z->minv[i] * z->p[i];
}
/* H = T + V with T = 1/2 p . (M^-1 p): diag_e_metric::T
 * (diag_e_metric.hpp:20-22), dense_e_metric::T (dense_e_metric.hpp:21-23) */
double orc_hamiltonian(const orc_point* z) {
  double t = 0.0;
  if (z->dense) {
    for (int i = 0; i < z->D; ++i) {
      const double* row = z->minv + (size_t)i * (size_t)z->D;
      double s = 0.0;
      for (int j = 0; j < z->D; ++j) s += row[j] * z->p[j];
      t += z->p[i] * s;
    }
  } else {
    for (int i = 0; i < z->D; ++i) t += z->p[i] * (z->minv[i] * z->p[i]);
  }
  return 0.5 * t + z->V;
}

/* Cholesky factor A = L L^T of a row-major symmetric matrix (lower L,
 * row-major; column by column, k ascending in every sum).  Returns 0, or
 * -1 when a pivot is not positive (A not positive definite).           */
int orc_cholesky(const double* A, int D, double* L) {
  memset(L, 0, sizeof(double) * (size_t)D * (size_t)D);
  for (int j = 0; j < D; ++j) {
    double s = A[(size_t)j * D + j];
    for (int k = 0; k < j; ++k) s -= L[(size_t)j * D + k] * L[(size_t)j * D + k];
    if (!(s > 0.0)) return -1;
    const double d = sqrt(s);
    L[(size_t)j * D + j] = d;
    for (int i = j + 1; i < D; ++i) {
      double t = A[(size_t)i * D + j];
      for (int k = 0; k < j; ++k) t -= L[(size_t)i * D + k] * L[(size_t)j * D + k];
      L[(size_t)i * D + j] = t / d;
    }
  }
  return 0;
}

/* dense_e_metric::sample_p (dense_e_metric.hpp:43-54): u_i ~ N(0,1), then
 * p = U^-1 u with U the upper Cholesky factor of M^-1 (M^-1 = U^T U,
 * U = L^T), by back substitution (j ascending in each row's sum).  p then
 * has covariance (U^T U)^-1 = M.  L is the factor of the current metric. */
static void solve_upper(const double* L, int D, const double* u, double* p) {
  for (int i = D - 1; i >= 0; --i) {
    double t = u[i];
    for (int j = i + 1; j < D; ++j) t -= L[(size_t)j * D + i] * p[j];
    p[i] = t / L[(size_t)i * D + i];
  }
}

/* base_hamiltonian::update_potential_gradient (base_hamiltonian.hpp:61-70).
 * Returns 0, or the fatal status of the model callback.                */
static int update_potential_gradient(orc_lpgrad_fn f, void* ctx,
                                     orc_point* z, int64_t* n_grad) {
  double lp = 0.0;
  double* g = vec(z->D);
  int st = f(ctx, z->q, z->D, &lp, g);
  if (n_grad) ++*n_grad;
  if (st == 0) {
    z->V = -lp;
    for (int i = 0; i < z->D; ++i) z->g[i] = -g[i];
  } else if (st == 1) {
    z->V = INFINITY;
    for (int i = 0; i < z->D; ++i) z->g[i] = -z->g[i];
  } else {
    free(g);
    return st;
  }
  free(g);
  return 0;
}

static int evolve(orc_lpgrad_fn f, void* ctx, orc_point* z, double eps,
                  int64_t* n_grad) {
  const int D = z->D;
  const double he = 0.5 * eps;
  for (int i = 0; i < D; ++i) z->p[i] -= he * z->g[i];
  if (z->dense) {
    double* ps = vec(D);
    dtau_dp(z, ps);
    for (int i = 0; i < D; ++i) z->q[i] += eps * ps[i];
    free(ps);
  } else {
    for (int i = 0; i < D; ++i) z->q[i] += eps * (z->minv[i] * z->p[i]);
  }
  int st = update_potential_gradient(f, ctx, z, n_grad);
  if (st) return st;
  for (int i = 0; i < D; ++i) z->p[i] -= he * z->g[i];
  return 0;
}

int orc_leapfrog(orc_lpgrad_fn f, void* ctx, orc_point* z, double eps) {
  return evolve(f, ctx, z, eps, NULL);
}

/* ---------------- sampler state ---------------- */
typedef struct sampler {
  orc_lpgrad_fn f;
  void* ctx;
  int D;
  orc_rng rng;
  double nom_eps, eps, jitter;
  int max_depth;
  double max_deltaH;
  int depth, n_leapfrog, divergent;
  double energy;
  orc_point z;
  double* minv;  /* D (diag) or D x D row-major (dense) */
  int dense;
  double* chol;  /* dense: Cholesky factor of minv, refreshed with it */
  int64_t n_grad;
  int fatal;
  /* stepsize_adaptation */
  double sa[8]; /* counter, s_bar, x_bar, mu, delta, gamma, kappa, t0 */
  /* windowed_adaptation */
  unsigned num_warmup, init_buffer, term_buffer, base_window;
  unsigned window_counter, next_window, window_size;
  /* welford estimator */
  double wn;
  double *wm, *wm2;
  int adapt_flag;
} sampler;

static int compute_criterion(const double* p_sharp_minus,
                             const double* p_sharp_plus, const double* rho,
                             int D) {
  return dot(p_sharp_plus, rho, D) > 0 && dot(p_sharp_minus, rho, D) > 0;
}

/* base_nuts::build_tree (base_nuts.hpp:247-352) */
static int build_tree_impl(sampler* s, int depth, orc_point* z_propose,
                           double* p_sharp_beg, double* p_sharp_end,
                           double* rho, double* p_beg, double* p_end,
                           double H0, double sign, int* n_leapfrog,
                           double* log_sum_weight, double* sum_metro_prob) {
  const int D = s->D;
  orc_point* z = &s->z;
  if (depth == 0) {
    int st = evolve(s->f, s->ctx, z, sign * s->eps, &s->n_grad);
    if (st) { s->fatal = st; return 0; }
    ++*n_leapfrog;
    double h = orc_hamiltonian(z);
    if (isnan(h)) h = INFINITY;
    if ((h - H0) > s->max_deltaH) s->divergent = 1;
    *log_sum_weight = log_sum_exp(*log_sum_weight, H0 - h);
    if (H0 - h > 0)
      *sum_metro_prob += 1;
    else
      *sum_metro_prob += exp(H0 - h);
    pt_copy(z_propose, z);
    dtau_dp(z, p_sharp_beg);
    memcpy(p_sharp_end, p_sharp_beg, sizeof(double) * D);
    for (int i = 0; i < D; ++i) rho[i] += z->p[i];
    memcpy(p_beg, z->p, sizeof(double) * D);
    memcpy(p_end, p_beg, sizeof(double) * D);
    return !s->divergent;
  }
  double log_sum_weight_init = NEG_INF;
  double* p_init_end = vec(D);
  double* p_sharp_init_end = vec(D);
  double* rho_init = vec(D);
  int valid_init = build_tree_impl(s, depth - 1, z_propose, p_sharp_beg,
                                   p_sharp_init_end, rho_init, p_beg,
                                   p_init_end, H0, sign, n_leapfrog,
                                   &log_sum_weight_init, sum_metro_prob);
  int result = 0;
  orc_point z_propose_final;
  pt_alloc(&z_propose_final, D, s->minv, s->dense);
  double* p_final_beg = vec(D);
  double* p_sharp_final_beg = vec(D);
  double* rho_final = vec(D);
  double* rho_subtree = vec(D);
  if (!valid_init) goto done;
  pt_copy(&z_propose_final, z);
  double log_sum_weight_final = NEG_INF;
  int valid_final = build_tree_impl(s, depth - 1, &z_propose_final,
                                    p_sharp_final_beg, p_sharp_end,
                                    rho_final, p_final_beg, p_end, H0, sign,
                                    n_leapfrog, &log_sum_weight_final,
                                    sum_metro_prob);
  if (!valid_final) goto done;
  double log_sum_weight_subtree =
      log_sum_exp(log_sum_weight_init, log_sum_weight_final);
  *log_sum_weight = log_sum_exp(*log_sum_weight, log_sum_weight_subtree);
  if (log_sum_weight_final > log_sum_weight_subtree) {
    pt_copy(z_propose, &z_propose_final);
  } else {
    double accept_prob = exp(log_sum_weight_final - log_sum_weight_subtree);
    if (orc_rng_uniform(&s->rng) < accept_prob)
      pt_copy(z_propose, &z_propose_final);
  }
  for (int i = 0; i < D; ++i) rho_subtree[i] = rho_init[i] + rho_final[i];
  for (int i = 0; i < D; ++i) rho[i] += rho_subtree[i];
  int persist = compute_criterion(p_sharp_beg, p_sharp_end, rho_subtree, D);
  for (int i = 0; i < D; ++i) rho_subtree[i] = rho_init[i] + p_final_beg[i];
  persist &= compute_criterion(p_sharp_beg, p_sharp_final_beg, rho_subtree, D);
  for (int i = 0; i < D; ++i) rho_subtree[i] = rho_final[i] + p_init_end[i];
  persist &= compute_criterion(p_sharp_init_end, p_sharp_end, rho_subtree, D);
  result = persist;
done:
  free(p_init_end); free(p_sharp_init_end); free(rho_init);
  pt_free(&z_propose_final);
  free(p_final_beg); free(p_sharp_final_beg); free(rho_final);
  free(rho_subtree);
  return result;
}

int orc_build_tree(orc_lpgrad_fn f, void* ctx, orc_rng* rng, double eps,
                   double max_deltaH, int depth, orc_point* z,
                   orc_point* z_propose, double* p_sharp_beg,
                   double* p_sharp_end, double* rho, double* p_beg,
                   double* p_end, double H0, double sign, int* n_leapfrog,
                   double* log_sum_weight, double* sum_metro_prob,
                   int* divergent) {
  sampler s;
  memset(&s, 0, sizeof(s));
  s.f = f; s.ctx = ctx; s.D = z->D; s.rng = *rng; s.eps = eps;
  s.max_deltaH = max_deltaH; s.minv = z->minv; s.dense = z->dense;
  s.z = *z; /* shares storage with the caller's point */
  s.divergent = *divergent;
  int v = build_tree_impl(&s, depth, z_propose, p_sharp_beg, p_sharp_end,
                          rho, p_beg, p_end, H0, sign, n_leapfrog,
                          log_sum_weight, sum_metro_prob);
  z->V = s.z.V;
  *rng = s.rng;
  *divergent = s.divergent;
  return s.fatal ? -s.fatal : v;
}

static void sample_p(sampler* s, orc_point* z) {
  if (s->dense) {
    double* u = vec(s->D);
    for (int i = 0; i < s->D; ++i) u[i] = orc_rng_normal(&s->rng);
    solve_upper(s->chol, s->D, u, z->p);
    free(u);
    return;
  }
  for (int i = 0; i < s->D; ++i)
    z->p[i] = orc_rng_normal(&s->rng) / sqrt(z->minv[i]);
}

/* base_hmc::sample_stepsize (base_hmc.hpp:195-200) */
static void sample_stepsize(sampler* s) {
  s->eps = s->nom_eps;
  if (s->jitter)
    s->eps *= 1.0 + s->jitter * (2.0 * orc_rng_uniform(&s->rng) - 1.0);
}

/* base_nuts::transition (base_nuts.hpp:78-204).  Writes the sample's
 * lp and accept_stat.  Returns 0 or a fatal status.                    */
static int nuts_transition(sampler* s, double* q_inout, double* out_lp,
                           double* out_accept) {
  const int D = s->D;
  orc_point* z = &s->z;
  sample_stepsize(s);
  memcpy(z->q, q_inout, sizeof(double) * D);
  sample_p(s, z);
  int st = update_potential_gradient(s->f, s->ctx, z, &s->n_grad);
  if (st) return st;

  orc_point z_fwd, z_bck, z_sample, z_propose;
  pt_alloc(&z_fwd, D, s->minv, s->dense); pt_copy(&z_fwd, z);
  pt_alloc(&z_bck, D, s->minv, s->dense); pt_copy(&z_bck, &z_fwd);
  pt_alloc(&z_sample, D, s->minv, s->dense); pt_copy(&z_sample, &z_fwd);
  pt_alloc(&z_propose, D, s->minv, s->dense); pt_copy(&z_propose, &z_fwd);

  double* p_fwd_fwd = vec(D); memcpy(p_fwd_fwd, z->p, sizeof(double) * D);
  double* p_sharp_fwd_fwd = vec(D); dtau_dp(z, p_sharp_fwd_fwd);
  double* p_fwd_bck = vec(D); memcpy(p_fwd_bck, z->p, sizeof(double) * D);
  double* p_sharp_fwd_bck = vec(D);
  memcpy(p_sharp_fwd_bck, p_sharp_fwd_fwd, sizeof(double) * D);
  double* p_bck_fwd = vec(D); memcpy(p_bck_fwd, z->p, sizeof(double) * D);
  double* p_sharp_bck_fwd = vec(D);
  memcpy(p_sharp_bck_fwd, p_sharp_fwd_fwd, sizeof(double) * D);
  double* p_bck_bck = vec(D); memcpy(p_bck_bck, z->p, sizeof(double) * D);
  double* p_sharp_bck_bck = vec(D);
  memcpy(p_sharp_bck_bck, p_sharp_fwd_fwd, sizeof(double) * D);
  double* rho = vec(D); memcpy(rho, z->p, sizeof(double) * D);
  double* rho_fwd = vec(D);
  double* rho_bck = vec(D);
  double* rho_ext = vec(D);

  double log_sum_weight = 0;
  double H0 = orc_hamiltonian(z);
  int n_leapfrog = 0;
  double sum_metro_prob = 0;
  s->depth = 0;
  s->divergent = 0;

  while (s->depth < s->max_depth) {
    memset(rho_fwd, 0, sizeof(double) * D);
    memset(rho_bck, 0, sizeof(double) * D);
    int valid_subtree = 0;
    double log_sum_weight_subtree = NEG_INF;
    if (orc_rng_uniform(&s->rng) > 0.5) {
      pt_copy(z, &z_fwd);
      memcpy(rho_bck, rho, sizeof(double) * D);
      memcpy(p_bck_fwd, p_fwd_fwd, sizeof(double) * D);
      memcpy(p_sharp_bck_fwd, p_sharp_fwd_fwd, sizeof(double) * D);
      valid_subtree = build_tree_impl(
          s, s->depth, &z_propose, p_sharp_fwd_bck, p_sharp_fwd_fwd, rho_fwd,
          p_fwd_bck, p_fwd_fwd, H0, 1, &n_leapfrog, &log_sum_weight_subtree,
          &sum_metro_prob);
      pt_copy(&z_fwd, z);
    } else {
      pt_copy(z, &z_bck);
      memcpy(rho_fwd, rho, sizeof(double) * D);
      memcpy(p_fwd_bck, p_bck_bck, sizeof(double) * D);
      memcpy(p_sharp_fwd_bck, p_sharp_bck_bck, sizeof(double) * D);
      valid_subtree = build_tree_impl(
          s, s->depth, &z_propose, p_sharp_bck_fwd, p_sharp_bck_bck, rho_bck,
          p_bck_fwd, p_bck_bck, H0, -1, &n_leapfrog, &log_sum_weight_subtree,
          &sum_metro_prob);
      pt_copy(&z_bck, z);
    }
    if (s->fatal) break;
    if (!valid_subtree) break;
    ++s->depth;
    if (log_sum_weight_subtree > log_sum_weight) {
      pt_copy(&z_sample, &z_propose);
    } else {
      double accept_prob = exp(log_sum_weight_subtree - log_sum_weight);
      if (orc_rng_uniform(&s->rng) < accept_prob)
        pt_copy(&z_sample, &z_propose);
    }
    log_sum_weight = log_sum_exp(log_sum_weight, log_sum_weight_subtree);
    for (int i = 0; i < D; ++i) rho[i] = rho_bck[i] + rho_fwd[i];
    int persist = compute_criterion(p_sharp_bck_bck, p_sharp_fwd_fwd, rho, D);
    for (int i = 0; i < D; ++i) rho_ext[i] = rho_bck[i] + p_fwd_bck[i];
    persist &= compute_criterion(p_sharp_bck_bck, p_sharp_fwd_bck, rho_ext, D);
    for (int i = 0; i < D; ++i) rho_ext[i] = rho_fwd[i] + p_bck_fwd[i];
    persist &= compute_criterion(p_sharp_bck_fwd, p_sharp_fwd_fwd, rho_ext, D);
    if (!persist) break;
  }
  int fatal = s->fatal;
  s->n_leapfrog = n_leapfrog;
  double accept_prob = sum_metro_prob / (double)n_leapfrog;
  pt_copy(z, &z_sample);
  s->energy = orc_hamiltonian(z);
  memcpy(q_inout, z->q, sizeof(double) * D);
  *out_lp = -z->V;
  *out_accept = accept_prob;

  pt_free(&z_fwd); pt_free(&z_bck); pt_free(&z_sample); pt_free(&z_propose);
  free(p_fwd_fwd); free(p_sharp_fwd_fwd); free(p_fwd_bck);
  free(p_sharp_fwd_bck); free(p_bck_fwd); free(p_sharp_bck_fwd);
  free(p_bck_bck); free(p_sharp_bck_bck); free(rho); free(rho_fwd);
  free(rho_bck); free(rho_ext);
  return fatal;
}

/* base_hmc::init_stepsize (base_hmc.hpp:93-161).  Returns 0, a fatal
 * model status, or -1 for the two runtime_error exits.                  */
static int init_stepsize(sampler* s) {
  const int D = s->D;
  orc_point* z = &s->z;
  if (D == 0) { s->nom_eps = NAN; return 0; }
  if (s->nom_eps == 0 || s->nom_eps > 1e7 || isnan(s->nom_eps)) return 0;
  orc_point z_init;
  pt_alloc(&z_init, D, s->minv, s->dense);
  pt_copy(&z_init, z);
  int st;
  sample_p(s, z);
  if ((st = update_potential_gradient(s->f, s->ctx, z, &s->n_grad))) goto fail;
  double H0 = orc_hamiltonian(z);
  if ((st = evolve(s->f, s->ctx, z, s->nom_eps, &s->n_grad))) goto fail;
  double h = orc_hamiltonian(z);
  if (isnan(h)) h = INFINITY;
  double delta_H = H0 - h;
  int direction = delta_H > log(0.8) ? 1 : -1;
  while (1) {
    pt_copy(z, &z_init);
    sample_p(s, z);
    if ((st = update_potential_gradient(s->f, s->ctx, z, &s->n_grad))) goto fail;
    double H0b = orc_hamiltonian(z);
    if ((st = evolve(s->f, s->ctx, z, s->nom_eps, &s->n_grad))) goto fail;
    double hb = orc_hamiltonian(z);
    if (isnan(hb)) hb = INFINITY;
    double dH = H0b - hb;
    if ((direction == 1) && !(dH > log(0.8)))
      break;
    else if ((direction == -1) && !(dH < log(0.8)))
      break;
    else
      s->nom_eps = direction == 1 ? 2.0 * s->nom_eps : 0.5 * s->nom_eps;
    if (s->nom_eps > 1e7) { st = -1; goto fail; }
    if (s->nom_eps == 0) { st = -1; goto fail; }
  }
  pt_copy(z, &z_init);
  pt_free(&z_init);
  return 0;
fail:
  pt_free(&z_init);
  return st;
}

/* stepsize_adaptation::learn_stepsize (stepsize_adaptation.hpp:55-71) */
void orc_learn_stepsize(double* sa, double* epsilon, double adapt_stat) {
  double* counter = &sa[0];
  double* s_bar = &sa[1];
  double* x_bar = &sa[2];
  double mu = sa[3], delta = sa[4], gamma = sa[5], kappa = sa[6], t0 = sa[7];
  ++*counter;
  adapt_stat = adapt_stat > 1 ? 1 : adapt_stat;
  const double eta = 1.0 / (*counter + t0);
  *s_bar = (1.0 - eta) * *s_bar + eta * (delta - adapt_stat);
  const double x = mu - *s_bar * sqrt(*counter) / gamma;
  const double x_eta = pow(*counter, -kappa);
  *x_bar = (1.0 - x_eta) * *x_bar + x_eta * x;
  *epsilon = exp(x);
}

/* windowed_adaptation (windowed_adaptation.hpp:23-110) */
static void wa_restart(sampler* s) {
  s->window_counter = 0;
  s->window_size = s->base_window;
  s->next_window = s->init_buffer + s->window_size - 1;
}
static void wa_set_params(sampler* s, unsigned num_warmup, unsigned init_buffer,
                          unsigned term_buffer, unsigned base_window) {
  if (num_warmup < 20) return;
  if (init_buffer + base_window + term_buffer > num_warmup) {
    s->num_warmup = num_warmup;
    s->init_buffer = (unsigned)(0.15 * num_warmup);
    s->term_buffer = (unsigned)(0.10 * num_warmup);
    s->base_window = num_warmup - (s->init_buffer + s->term_buffer);
    return;
  }
  s->num_warmup = num_warmup;
  s->init_buffer = init_buffer;
  s->term_buffer = term_buffer;
  s->base_window = base_window;
  wa_restart(s);
}
static int wa_adaptation_window(const sampler* s) {
  return (s->window_counter >= s->init_buffer) &&
         (s->window_counter < s->num_warmup - s->term_buffer) &&
         (s->window_counter != s->num_warmup);
}
static int wa_end_adaptation_window(const sampler* s) {
  return (s->window_counter == s->next_window) &&
         (s->window_counter != s->num_warmup);
}
static void wa_compute_next_window(sampler* s) {
  if (s->next_window == s->num_warmup - s->term_buffer - 1) return;
  s->window_size *= 2;
  s->next_window = s->window_counter + s->window_size;
  if (s->next_window == s->num_warmup - s->term_buffer - 1) return;
  unsigned next_window_boundary = s->next_window + 2 * s->window_size;
  if (next_window_boundary >= s->num_warmup - s->term_buffer)
    s->next_window = s->num_warmup - s->term_buffer - 1;
}

/* var_adaptation::learn_variance (var_adaptation.hpp:17-46) with the
 * Welford running variance.  Returns 1 when the metric was updated, 0 if
 * not, -1 on non-finite metric (runtime_error in the reference).       */
static int learn_variance(sampler* s, double* var, const double* q) {
  const int D = s->D;
  if (wa_adaptation_window(s)) {
    s->wn += 1.0;
    for (int i = 0; i < D; ++i) {
      double delta = q[i] - s->wm[i];
      s->wm[i] += delta / s->wn;
      s->wm2[i] += (q[i] - s->wm[i]) * delta;
    }
  }
  if (wa_end_adaptation_window(s)) {
    wa_compute_next_window(s);
    if (s->wn > 1)
      for (int i = 0; i < D; ++i) var[i] = s->wm2[i] / (s->wn - 1.0);
    double n = s->wn;
    int finite = 1;
    for (int i = 0; i < D; ++i) {
      var[i] = (n / (n + 5.0)) * var[i] + 1e-3 * (5.0 / (n + 5.0)) * 1.0;
      if (!isfinite(var[i])) finite = 0;
    }
    if (!finite) return -1;
    s->wn = 0;
    memset(s->wm, 0, sizeof(double) * D);
    memset(s->wm2, 0, sizeof(double) * D);
    ++s->window_counter;
    return 1;
  }
  ++s->window_counter;
  return 0;
}

/* var_adaptation_test.cpp:5-27: a variance window fed only zero draws.
 * Constructs the window state as the reference's constructor and
 * set_window_params do, calls learn_variance n_learn times with q = 0 and
 * returns the metric (started at 0) in var. */
int orc_var_window_zero(int D, int n_learn, unsigned num_warmup, unsigned init_buffer,
                        unsigned term_buffer, unsigned base_window, double* var) {
  sampler s;
  memset(&s, 0, sizeof(s));
  s.D = D;
  s.wm = vec(D);
  s.wm2 = vec(D);
  double* q = vec(D);
  wa_restart(&s);
  wa_set_params(&s, num_warmup, init_buffer, term_buffer, base_window);
  for (int i = 0; i < D; ++i) var[i] = 0.0;
  int rc = 0;
  for (int i = 0; i < n_learn && rc == 0; ++i)
    if (learn_variance(&s, var, q) < 0) rc = -1;
  free(q);
  free(s.wm);
  free(s.wm2);
  return rc;
}

/* covar_adaptation::learn_covariance (covar_adaptation.hpp:17-47) with the
 * Welford running covariance of Stan Math's welford_covar_estimator
 * (absent from the tree; restated from its definition: delta = q - m,
 * m += delta / n, m2 += (q - m) delta^T, covariance m2 / (n - 1)).  The
 * regularised estimate n/(n+5) C + 1e-3 5/(n+5) I replaces the metric
 * (row-major) and its Cholesky factor.  Returns 1 when the metric was
 * updated, 0 if not, -1 on a non-finite metric.                          */
static int learn_covariance(sampler* s, double* covar, const double* q) {
  const int D = s->D;
  if (wa_adaptation_window(s)) {
    s->wn += 1.0;
    double* delta = vec(D);
    for (int i = 0; i < D; ++i) delta[i] = q[i] - s->wm[i];
    for (int i = 0; i < D; ++i) s->wm[i] += delta[i] / s->wn;
    for (int i = 0; i < D; ++i) {
      const double d = q[i] - s->wm[i];
      for (int j = 0; j < D; ++j) s->wm2[(size_t)i * D + j] += d * delta[j];
    }
    free(delta);
  }
  if (wa_end_adaptation_window(s)) {
    wa_compute_next_window(s);
    const size_t DD = (size_t)D * (size_t)D;
    if (s->wn > 1)
      for (size_t k = 0; k < DD; ++k) covar[k] = s->wm2[k] / (s->wn - 1.0);
    double n = s->wn;
    const double a = n / (n + 5.0), b = 1e-3 * (5.0 / (n + 5.0));
    int finite = 1;
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) {
        double* c = &covar[(size_t)i * D + j];
        *c = a * *c + b * (i == j ? 1.0 : 0.0);
        if (!isfinite(*c)) finite = 0;
      }
    if (!finite) return -1;
    if (orc_cholesky(covar, D, s->chol) != 0) return -1;
    s->wn = 0;
    memset(s->wm, 0, sizeof(double) * D);
    memset(s->wm2, 0, sizeof(double) * DD);
    ++s->window_counter;
    return 1;
  }
  ++s->window_counter;
  return 0;
}

/* adapt_diag_e_nuts::transition (adapt_diag_e_nuts.hpp:25-43) and
 * adapt_dense_e_nuts::transition (adapt_dense_e_nuts.hpp:25-43) */
static int adaptive_transition(sampler* s, double* q, double* lp,
                               double* accept) {
  int st = nuts_transition(s, q, lp, accept);
  if (st) return st;
  if (s->adapt_flag) {
    orc_learn_stepsize(s->sa, &s->nom_eps, *accept);
    int update = s->dense ? learn_covariance(s, s->minv, s->z.q)
                          : learn_variance(s, s->minv, s->z.q);
    if (update < 0) return -1;
    if (update) {
      st = init_stepsize(s);
      if (st) return st;
      s->sa[3] = log(10 * s->nom_eps);
      s->sa[0] = 0; s->sa[1] = 0; s->sa[2] = 0;
    }
  }
  return 0;
}

static void write_row(const sampler* s, const double* q, double lp,
                      double accept, double* row) {
  row[0] = lp;
  row[1] = accept;
  row[2] = s->eps;
  row[3] = s->depth;
  row[4] = s->n_leapfrog;
  row[5] = s->divergent;
  row[6] = s->energy;
  memcpy(row + 7, q, sizeof(double) * s->D);
}

/* The services: hmc_nuts_diag_e_adapt (hmc_nuts_diag_e_adapt.hpp:57-112),
 * hmc_nuts_dense_e_adapt (hmc_nuts_dense_e_adapt.hpp:57-112) through
 * run_adaptive_sampler, and hmc_nuts_dense_e (hmc_nuts_dense_e.hpp:48-93) through
 * run_sampler (no step-size heuristic, no adaptation). */
static int run_service(orc_lpgrad_fn f, void* ctx, int D, const orc_nuts_args* a,
                       int dense, int adapt, double* out_draws, int64_t* out_rows,
                       double* out_stepsize, double* out_inv_metric, int64_t* n_grad) {
  sampler s;
  memset(&s, 0, sizeof(s));
  s.f = f; s.ctx = ctx; s.D = D; s.dense = dense;
  orc_rng_init(&s.rng, a->seed, a->chain);
  *out_rows = 0;
  const size_t DD = dense ? (size_t)D * (size_t)D : (size_t)D;
  double* q = vec(D);
  double* gbuf = vec(D);
  /* util::initialize (initialize.hpp:68-227) */
  int ok = 0;
  int zero_init = a->init_radius == 0.0;
  int max_tries = (a->init || zero_init) ? 1 : 100;
  for (int t = 0; t < max_tries && !ok; ++t) {
    if (a->init) {
      memcpy(q, a->init, sizeof(double) * D);
    } else {
      for (int i = 0; i < D; ++i)
        q[i] = zero_init ? 0.0
                         : -a->init_radius +
                               orc_rng_uniform(&s.rng) * 2.0 * a->init_radius;
    }
    double lp;
    int st = f(ctx, q, D, &lp, gbuf);
    ++s.n_grad;
    if (st == 1) continue;
    if (st != 0) { free(q); free(gbuf); return 70; }
    if (!isfinite(lp)) continue;
    double gs = 0.0;
    for (int i = 0; i < D; ++i) gs += gbuf[i];
    if (isfinite(gs)) ok = 1;
  }
  free(gbuf);
  if (!ok) { free(q); return 78; }
  s.minv = (double*)calloc(DD + 1, sizeof(double));
  if (dense) {
    /* read_dense_inv_metric + validate_dense_inv_metric: symmetric (to
     * 1e-8), finite and positive definite, else CONFIG */
    s.chol = (double*)calloc(DD + 1, sizeof(double));
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j)
        s.minv[(size_t)i * D + j] =
            a->init_inv_metric ? a->init_inv_metric[(size_t)i * D + j] : (i == j ? 1.0 : 0.0);
    int bad = 0;
    for (int i = 0; i < D && !bad; ++i)
      for (int j = 0; j < D; ++j) {
        const double x = s.minv[(size_t)i * D + j], y = s.minv[(size_t)j * D + i];
        if (!isfinite(x) || fabs(x - y) > 1e-8) { bad = 1; break; }
      }
    if (bad || orc_cholesky(s.minv, D, s.chol) != 0) {
      free(q); free(s.minv); free(s.chol); return 78;
    }
  } else {
    for (int i = 0; i < D; ++i) {
      s.minv[i] = a->init_inv_metric ? a->init_inv_metric[i] : 1.0;
      if (!(s.minv[i] > 0) || !isfinite(s.minv[i])) {
        free(q); free(s.minv); return 78;
      }
    }
  }
  pt_alloc(&s.z, D, s.minv, s.dense);
  s.wm = vec(D);
  s.wm2 = (double*)calloc(DD + 1, sizeof(double));
  s.nom_eps = 0.1; /* base_hmc default nom_epsilon_ */
  if (a->stepsize > 0) s.nom_eps = a->stepsize;
  if (a->stepsize_jitter > 0 && a->stepsize_jitter < 1) s.jitter = a->stepsize_jitter;
  s.max_depth = 5;
  if (a->max_depth > 0) s.max_depth = a->max_depth;
  s.max_deltaH = 1000;
  /* stepsize_adaptation defaults then setters (hmc_nuts_diag_e_adapt.hpp:93-97) */
  s.sa[3] = 0.5; s.sa[4] = 0.5; s.sa[5] = 0.05; s.sa[6] = 0.75; s.sa[7] = 10;
  s.sa[3] = log(10 * a->stepsize);
  if (a->delta > 0 && a->delta < 1) s.sa[4] = a->delta;
  if (a->gamma > 0) s.sa[5] = a->gamma;
  if (a->kappa > 0) s.sa[6] = a->kappa;
  if (a->t0 > 0) s.sa[7] = a->t0;
  wa_restart(&s);
  wa_set_params(&s, (unsigned)a->num_warmup, a->init_buffer, a->term_buffer,
                a->window);

  int rc = 0;
  const int ncol = 7 + D;
  int64_t rows = 0;
  double lp = 0, acc = 0;
  memcpy(s.z.q, q, sizeof(double) * D);
  /* run_adaptive_sampler (run_adaptive_sampler.hpp:46-103): the step-size
   * heuristic first; run_sampler (run_sampler.hpp:39-85) has none */
  s.adapt_flag = adapt;
  if (adapt && init_stepsize(&s) != 0) goto out; /* logged and returned in the reference */
  for (int m = 0; m < a->num_warmup; ++m) {
    int st = adaptive_transition(&s, q, &lp, &acc);
    if (st) { rc = 70; goto out; }
    if (a->save_warmup && (m % a->num_thin) == 0)
      write_row(&s, q, lp, acc, out_draws + (rows++) * ncol);
  }
  if (adapt) {
    s.adapt_flag = 0;
    s.nom_eps = exp(s.sa[2]); /* complete_adaptation */
  }
  for (int m = 0; m < a->num_samples; ++m) {
    int st = adaptive_transition(&s, q, &lp, &acc);
    if (st) { rc = 70; goto out; }
    if ((m % a->num_thin) == 0)
      write_row(&s, q, lp, acc, out_draws + (rows++) * ncol);
  }
  *out_rows = rows;
out:
  if (out_stepsize) *out_stepsize = s.nom_eps;
  if (out_inv_metric) memcpy(out_inv_metric, s.minv, sizeof(double) * DD);
  if (n_grad) *n_grad = s.n_grad;
  pt_free(&s.z);
  free(s.wm); free(s.wm2); free(s.minv); free(s.chol); free(q);
  return rc;
}

int orc_hmc_nuts_diag_e_adapt(orc_lpgrad_fn f, void* ctx, int D, const orc_nuts_args* a,
                              double* out_draws, int64_t* out_rows, double* out_stepsize,
                              double* out_inv_metric, int64_t* n_grad) {
  return run_service(f, ctx, D, a, 0, 1, out_draws, out_rows, out_stepsize, out_inv_metric,
                     n_grad);
}
int orc_hmc_nuts_diag_e(orc_lpgrad_fn f, void* ctx, int D, const orc_nuts_args* a,
                        double* out_draws, int64_t* out_rows, int64_t* n_grad) {
  return run_service(f, ctx, D, a, 0, 0, out_draws, out_rows, NULL, NULL, n_grad);
}
int orc_hmc_nuts_dense_e_adapt(orc_lpgrad_fn f, void* ctx, int D, const orc_nuts_args* a,
                               double* out_draws, int64_t* out_rows, double* out_stepsize,
                               double* out_inv_metric, int64_t* n_grad) {
  return run_service(f, ctx, D, a, 1, 1, out_draws, out_rows, out_stepsize, out_inv_metric,
                     n_grad);
}
int orc_hmc_nuts_dense_e(orc_lpgrad_fn f, void* ctx, int D, const orc_nuts_args* a,
                         double* out_draws, int64_t* out_rows, int64_t* n_grad) {
  return run_service(f, ctx, D, a, 1, 0, out_draws, out_rows, NULL, NULL, n_grad);
}
