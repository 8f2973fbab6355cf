/*
 * stanx.h -- C-ABI of the MI355X (gfx950) GLM log-density/gradient engine
 * and the NUTS sampler that drives it.
 *
 * The entry points replace, on the GPU, the calls the reference makes at
 * the model boundary (paths relative to the stan-dev/stan 2.35.0 tree):
 *
 *   sx_glm_logp_grad       stan::model::log_prob_grad<propto,jacobian>
 *                          (src/stan/model/log_prob_grad.hpp:28-50, 68-89) and
 *                          stan::model::gradient (src/stan/model/gradient.hpp:14-35)
 *                          for a model whose log_prob is a
 *                          bernoulli_logit_glm / normal_id_glm likelihood
 *                          (Stan Math, reached through
 *                          src/stan/model/model_functional.hpp:21) plus priors.
 *   sx_glm_logp            model_base::log_prob / log_prob_propto (double)
 *                          (src/stan/model/model_base.hpp:191-302, 404-544;
 *                          src/stan/model/log_prob_propto.hpp:32-52)
 *   sx_glm_num_params      prob_grad::num_params_r (src/stan/model/prob_grad.hpp:57)
 *   sx_glm_logp_grad_batch many chains at once; the reference runs chains on
 *                          separate threads instead
 *                          (src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:387-401)
 *   sx_hmc_nuts_diag_e_adapt
 *                          stan::services::sample::hmc_nuts_diag_e_adapt
 *                          (src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:57-112),
 *                          same arguments and return codes, draws returned
 *                          as a table instead of through writer callbacks.
 *   sx_hmc_nuts_diag_e_adapt_cb
 *                          the same service for a model given as a C callback.
 *   sx_hmc_nuts_diag_e, sx_hmc_nuts_diag_e_cb
 *                          stan::services::sample::hmc_nuts_diag_e
 *                          (src/stan/services/sample/hmc_nuts_diag_e.hpp:49-91):
 *                          the same transition without adaptation.
 *   sx_hmc_nuts_dense_e_adapt(_cb, _csv, _cb_csv), sx_hmc_nuts_dense_e(_cb)
 *                          the dense_e services
 *                          (src/stan/services/sample/hmc_nuts_dense_e_adapt.hpp:57-112,
 *                          src/stan/services/sample/hmc_nuts_dense_e.hpp:48-93).
 *
 * Conventions
 *   - Host buffers are caller-owned.  X (row-major N x K doubles) and y are
 *     copied to the device once, at create; the context owns device memory.
 *   - Every call is synchronous on the context's own HIP stream.  Calls on
 *     different contexts may run concurrently from different threads.
 *   - Return value: SX_OK, or an SX_E* status; sx_last_error() holds a
 *     thread-local message.  SX_EDOMAIN is the reference's std::domain_error
 *     (proposal rejected, src/stan/mcmc/hmc/hamiltonians/base_hamiltonian.hpp:61-70);
 *     any other error is fatal (std::runtime_error).
 *   - Service entry points return the reference's error codes
 *     (src/stan/services/error_codes.hpp:7-18): 0 OK, 70 SOFTWARE, 78 CONFIG.
 */
#ifndef STANX_H
#define STANX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SX_OK 0
#define SX_EDOMAIN 1
#define SX_EINVAL 2
#define SX_EDEVICE 3
#define SX_ENOMEM 4
#define SX_ETIMEOUT 5

/* families: theta layouts (unconstrained)
 *   SX_LOGIT      : [alpha, beta_1..beta_K]                      D = K+1
 *   SX_NORMAL     : [alpha, beta_1..beta_K, u], sigma = exp(u)   D = K+2
 *   SX_HIER_LOGIT : [alpha, beta_1..K, v_0..v_K, z_{1,0..K} .. z_{J,0..K}],
 *                   tau_k = exp(v_k)                     D = 2K+2+J(K+1)  */
#define SX_LOGIT 0
#define SX_NORMAL 1
#define SX_HIER_LOGIT 2

/* storage/arithmetic precision of the bulk X pass */
#define SX_FP64 0   /* FP64 X, FP64 arithmetic: the plain drop-in           */
#define SX_FP32 1   /* SX_HIER_LOGIT: FP64 X for single evaluations, plus X
                        in FP32 for batched evaluations (FP32 products and
                        chunk sums, FP64 sums over chunks and groups)   */
#define SX_BF16X2 2 /* FP64 X for single evaluations, plus X as BF16 hi+lo
                        for batched MFMA evaluations (K <= 512)        */

/* service error codes (src/stan/services/error_codes.hpp:7-18) */
#define SX_CODE_OK 0
#define SX_CODE_SOFTWARE 70
#define SX_CODE_CONFIG 78

typedef struct sx_glm sx_glm; /* opaque */

typedef struct sx_prior {
  double alpha_sd; /* alpha ~ normal(0, alpha_sd); <= 0 means flat   */
  double beta_sd;  /* beta  ~ normal(0, beta_sd)                     */
  double sigma_sd; /* sigma ~ half-normal(0, sigma_sd)  (SX_NORMAL)  */
  double tau_sd;   /* tau   ~ half-normal(0, tau_sd)    (SX_HIER_LOGIT) */
  double sigma_mu; /* location of the sigma prior: sigma ~ normal(sigma_mu,
                      sigma_sd) on sigma > 0 (0: the half-normal above)  */
} sx_prior;

/* y: const int32_t* with values in {0,1} for SX_LOGIT / SX_HIER_LOGIT,
 * const double* for SX_NORMAL.  group_offsets: J+1 CSR row offsets for
 * SX_HIER_LOGIT (rows of one group are contiguous), NULL otherwise.
 * Replaces the data constructor of a generated model (the model owns its
 * data: src/stan/model/model_base.hpp:36-56).                          */
int sx_glm_create(sx_glm** out, int family, const double* X, const void* y,
                  int64_t N, int64_t K, const int64_t* group_offsets,
                  int64_t J, int precision, const sx_prior* prior,
                  int device);
/* Kernel tuning of the bulk X pass.  Zero fields mean "choose":
 *   lanes_per_row  lanes sharing one row of X (power of two <= 64); fixes
 *                  the tiled layout of X in HBM
 *   slots_cap      compiled register capacity per lane (multiple of 4, <= 32)
 *   prefetch       1: double-buffer tiles in registers; 0: single buffer;
 *                  -1: choose
 *   blocks         workgroups per launch
 *   autotune       1: time the candidate configurations on the device at
 *                  create and keep the fastest (default for N*K >= 2^22):
 *                  the FP64 pass's layout and grid, the MFMA pass's kernel
 *                  variant and row split, the hierarchical chunk size
 *   hier_chunk     SX_HIER_LOGIT: rows per chunk at most (a multiple of 32,
 *                  <= 4096; 0: autotuned, else 256)
 *   hier_kernel    SX_HIER_LOGIT with SX_FP32: 1 = FP32 MFMA products,
 *                  2 = products of exact three-term BF16 splits (K <= 31;
 *                  above, 1), 0 = autotuned, else 1
 * No counterpart in the reference (its CPU GLM has no tunables).         */
typedef struct sx_tuning {
  int lanes_per_row;
  int slots_cap;
  int prefetch;
  int blocks;
  int autotune;
  int hier_chunk;
  int hier_kernel;
} sx_tuning;

/* sx_glm_create with explicit tuning (NULL = defaults). */
int sx_glm_create_ex(sx_glm** out, int family, const double* X, const void* y,
                     int64_t N, int64_t K, const int64_t* group_offsets,
                     int64_t J, int precision, const sx_prior* prior,
                     int device, const sx_tuning* tuning);
/* A context from a Stan JSON data file, read as a generated model reads its
 * var_context (src/stan/io/json/json_data.hpp:42-73): X (N x K matrix), y
 * (or Y); N and K when present must agree; the hierarchical model takes a
 * 1-based group index per row (group or ids) and J (else the largest
 * index), and its rows are grouped by a stable sort.                      */
int sx_glm_create_json(sx_glm** out, const char* path, int family,
                       int precision, const sx_prior* prior, int device);
/* One variable of a Stan JSON data file: values column-major (first index
 * fastest, the var_context layout), up to cap of them; dims up to
 * max_dims; host only.                                                    */
int sx_json_var(const char* path, const char* name, double* vals,
                int64_t cap, int64_t* dims, int max_dims, int* ndims,
                int64_t* size);
/* Text log of the autotuner's measurements and the chosen configuration. */
const char* sx_glm_tuning_log(const sx_glm* m);
void sx_glm_destroy(sx_glm* m);
int sx_glm_num_params(const sx_glm* m, int64_t* D);
/* data sizes the context was created with (any pointer may be NULL);
 * a generated model knows these from its data block                   */
int sx_glm_dims(const sx_glm* m, int64_t* N, int64_t* K, int64_t* J);
/* family the context was created with (SX_LOGIT, ...), or -1 */
int sx_glm_family(const sx_glm* m);
const char* sx_last_error(void);

/* log density (propto / jacobian as in log_prob<propto, jacobian>) and
 * its gradient at theta (D doubles, host).  grad may be NULL.          */
int sx_glm_logp_grad(sx_glm* m, const double* theta, int propto,
                     int jacobian, double* lp, double* grad);
int sx_glm_logp(sx_glm* m, const double* theta, int propto, int jacobian,
                double* lp);

/* C chains at once: theta is C x D (row-major), lp C, grad C x D.  On a
 * context created with SX_BF16X2 all chains are evaluated together by the
 * MFMA pass (X as BF16 hi+lo, FP32 accumulation, FP64 cross-block sums);
 * SX_HIER_LOGIT evaluates all chains in one pass (FP32 with SX_FP32, else
 * FP64, each chain's result bitwise its single evaluation); otherwise
 * chain by chain in FP64.  A chain whose density cannot be
 * evaluated gets lp = -inf and the call returns SX_EDOMAIN.             */
int sx_glm_logp_grad_batch(sx_glm* m, const double* theta, int64_t C,
                           int propto, int jacobian, double* lp,
                           double* grad);

/* Device-resident timing of the gradient evaluation: `iters` evaluations
 * at theta (copied to the device once) after `warmup` untimed ones,
 * bracketed by HIP events on the context's stream.  ms_per_eval: mean
 * time of a full evaluation (bulk X pass + cross-block reduction).
 * ms_kernel: mean duration of the bulk X pass alone, from a second loop
 * that launches only that kernel.  C > 1 times the batched evaluation of
 * sx_glm_logp_grad_batch (SX_BF16X2 contexts and SX_HIER_LOGIT); then
 * ms_kernel = ms_per_eval.                                              */
int sx_glm_time_logp_grad(sx_glm* m, const double* theta, int64_t C,
                          int iters, int warmup, double* ms_per_eval,
                          double* ms_kernel);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  One log
 * density + gradient evaluation at theta, already on the device (D
 * doubles), into out on the device (D + 2 doubles: lp, gradient, then a
 * status word that is nonzero when the density cannot be evaluated, the
 * SX_EDOMAIN case of sx_glm_logp_grad), enqueued on `stream` (a
 * hipStream_t; NULL: the context's stream) without waiting.  The same
 * kernels as sx_glm_logp_grad, with their partial sums in a per-GPU scratch
 * array of SX_NOALLOC_SCRATCH_BYTES that the library holds until the
 * process ends: no device memory is allocated or freed, no copy or memset
 * is issued.  A context whose partial sums need more (FP64 pass: blocks x
 * (K + 3) doubles; hierarchical: (K + 3) per row chunk + (2K + 5) per 8
 * groups) gets SX_EINVAL.  Calls that overlap on one GPU share the array
 * and race; the array keeps the last call's contents.                   */
#define SX_NOALLOC_SCRATCH_BYTES (8 * 1024 * 1024)
int sx_glm_logp_grad_noalloc(sx_glm* m, const double* d_theta, int propto,
                             int jacobian, double* d_out, void* stream);

/* Tuning knobs of the current kernel configuration (for reports):
 * lanes per row, slots per lane, blocks, threads per block.           */
int sx_glm_kernel_config(const sx_glm* m, int* lanes_per_row,
                         int* slots_per_lane, int* blocks, int* threads,
                         int* prefetch);

/* ---------------- sampler service ---------------- */
/* Arguments of hmc_nuts_diag_e_adapt in the reference's order
 * (src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:58-68).          */
typedef struct sx_nuts_config {
  uint32_t random_seed;
  uint32_t chain;
  double init_radius;
  int32_t num_warmup;
  int32_t num_samples;
  int32_t num_thin;
  int32_t save_warmup;
  int32_t refresh;
  double stepsize;
  double stepsize_jitter;
  int32_t max_depth;
  double delta;
  double gamma;
  double kappa;
  double t0;
  uint32_t init_buffer;
  uint32_t term_buffer;
  uint32_t window;
} sx_nuts_config;

/* Outputs of one run.  draws: rows x (7 + D) doubles with columns
 *   lp__, accept_stat__, stepsize__, treedepth__, n_leapfrog__,
 *   divergent__, energy__, then the D constrained parameters of
 *   model.write_array (sigma = exp(u), tau = exp(v); the others as is)
 * (src/stan/mcmc/base_nuts.hpp:206-220, src/stan/mcmc/sample.hpp:42-45).
 * Capacity must be (save_warmup ? num_warmup : 0) + num_samples rows
 * (before thinning).                                                   */
typedef struct sx_nuts_result {
  double* draws;
  int64_t rows;
  double stepsize;      /* adapted step size */
  double* inv_metric;   /* D doubles: adapted diagonal inverse metric */
  int64_t n_grad;       /* gradient evaluations, incl. init and heuristics */
  double warmup_seconds;
  double sampling_seconds;
} sx_nuts_result;

/* init / init_inv_metric: D doubles each, or NULL (random inits within
 * init_radius; unit metric).                                           */
int sx_hmc_nuts_diag_e_adapt(sx_glm* m, const sx_nuts_config* cfg,
                             const double* init,
                             const double* init_inv_metric,
                             sx_nuts_result* out);

/* Model callback: value and gradient of the log density (propto=true,
 * jacobian=true) at q.  Return 0, 1 for a domain error, anything else for
 * a fatal error.                                                        */
typedef int (*sx_lpgrad_fn)(void* ctx, const double* q, int D, double* lp,
                            double* grad);
int sx_hmc_nuts_diag_e_adapt_cb(sx_lpgrad_fn f, void* ctx, int D,
                                const sx_nuts_config* cfg,
                                const double* init,
                                const double* init_inv_metric,
                                sx_nuts_result* out);

/* The adaptive service writing its sample output as a Stan CSV file, as
 * the reference's services do through a stream_writer with comment prefix
 * "# " (src/stan/callbacks/stream_writer.hpp:14-93): a configuration block
 * in CmdStan's layout, the header (lp__, accept_stat__, ..., parameter
 * names), warm-up rows when saved, "Adaptation terminated", the step size
 * and inverse metric lines, the draws, and the timing block
 * (src/stan/services/util/mcmc_writer.hpp:91-238).  sig_figs > 0 sets the
 * significant digits of the values (the stream default otherwise).  out
 * receives the same rows as the table entry points.                      */
int sx_hmc_nuts_diag_e_adapt_csv(sx_glm* m, const sx_nuts_config* cfg,
                                 const double* init,
                                 const double* init_inv_metric,
                                 const char* csv_path, int sig_figs,
                                 sx_nuts_result* out);
int sx_hmc_nuts_diag_e_adapt_cb_csv(sx_lpgrad_fn f, void* ctx, int D,
                                    const sx_nuts_config* cfg,
                                    const double* init,
                                    const double* init_inv_metric,
                                    const char* csv_path, int sig_figs,
                                    sx_nuts_result* out);

/* The services with all three of the reference's file outputs
 * (src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:57-112: sample_writer,
 * diagnostic_writer, metric_writer; hmc_nuts_diag_e.hpp:49-91 has no
 * metric_writer):
 *   sample_csv      the Stan CSV of the _csv entry points;
 *   diagnostic_csv  per saved draw lp__ .. energy__, then the unconstrained
 *                   position, momentum (p_*) and potential gradient (g_*)
 *                   of the selected point (mcmc_writer.hpp:157-190,
 *                   ps_point.hpp:25-44), after the same configuration block;
 *   metric_json     one JSON record {"stepsize", "metric_type", "inv_metric"}
 *                   written when warm-up ends (base_hmc.hpp:70-76).
 * Any path may be NULL (that output is not written).                     */
typedef struct sx_output_files {
  const char* sample_csv;
  const char* diagnostic_csv;
  const char* metric_json;
  int sig_figs; /* > 0: significant digits of the CSV values */
} sx_output_files;
int sx_hmc_nuts_diag_e_adapt_files(sx_glm* m, const sx_nuts_config* cfg,
                                   const double* init, const double* init_inv_metric,
                                   const sx_output_files* files, sx_nuts_result* out);
int sx_hmc_nuts_diag_e_adapt_cb_files(sx_lpgrad_fn f, void* ctx, int D,
                                      const sx_nuts_config* cfg, const double* init,
                                      const double* init_inv_metric,
                                      const sx_output_files* files, sx_nuts_result* out);
int sx_hmc_nuts_diag_e_files(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                             const double* init_inv_metric, const sx_output_files* files,
                             sx_nuts_result* out);

/* hmc_nuts_diag_e (src/stan/services/sample/hmc_nuts_diag_e.hpp:49-91)
 * through run_sampler (src/stan/services/util/run_sampler.hpp:39-85): the
 * step size cfg->stepsize (with jitter) and the metric init_inv_metric (or
 * unit) stay fixed; warm-up iterations are transitions without
 * adaptation and no step-size heuristic runs.  delta, gamma, kappa, t0
 * and the window parameters of cfg are ignored; out->stepsize and
 * out->inv_metric return the fixed values.                              */
int sx_hmc_nuts_diag_e(sx_glm* m, const sx_nuts_config* cfg,
                       const double* init, const double* init_inv_metric,
                       sx_nuts_result* out);
int sx_hmc_nuts_diag_e_cb(sx_lpgrad_fn f, void* ctx, int D,
                          const sx_nuts_config* cfg, const double* init,
                          const double* init_inv_metric, sx_nuts_result* out);

/* The dense-metric services:
 *   sx_hmc_nuts_dense_e_adapt  stan::services::sample::hmc_nuts_dense_e_adapt
 *                              (src/stan/services/sample/hmc_nuts_dense_e_adapt.hpp:57-112)
 *   sx_hmc_nuts_dense_e        stan::services::sample::hmc_nuts_dense_e
 *                              (src/stan/services/sample/hmc_nuts_dense_e.hpp:48-93)
 * The same arguments as the diag_e entry points, except that
 * init_inv_metric is a D x D matrix (row-major; NULL = identity, as
 * util::create_unit_e_dense_inv_metric) that must be finite, symmetric to
 * 1e-8 and positive definite (util::validate_dense_inv_metric; otherwise
 * SX_CODE_CONFIG), and that out->inv_metric must hold D x D doubles: the
 * metric adapted by windowed covariance estimation
 * (src/stan/mcmc/covar_adaptation.hpp:17-47), or the fixed one.         */
int sx_hmc_nuts_dense_e_adapt(sx_glm* m, const sx_nuts_config* cfg,
                              const double* init, const double* init_inv_metric,
                              sx_nuts_result* out);
int sx_hmc_nuts_dense_e_adapt_cb(sx_lpgrad_fn f, void* ctx, int D,
                                 const sx_nuts_config* cfg, const double* init,
                                 const double* init_inv_metric, sx_nuts_result* out);
int sx_hmc_nuts_dense_e_adapt_csv(sx_glm* m, const sx_nuts_config* cfg,
                                  const double* init, const double* init_inv_metric,
                                  const char* csv_path, int sig_figs,
                                  sx_nuts_result* out);
int sx_hmc_nuts_dense_e_adapt_cb_csv(sx_lpgrad_fn f, void* ctx, int D,
                                     const sx_nuts_config* cfg, const double* init,
                                     const double* init_inv_metric,
                                     const char* csv_path, int sig_figs,
                                     sx_nuts_result* out);
int sx_hmc_nuts_dense_e(sx_glm* m, const sx_nuts_config* cfg,
                        const double* init, const double* init_inv_metric,
                        sx_nuts_result* out);
int sx_hmc_nuts_dense_e_cb(sx_lpgrad_fn f, void* ctx, int D,
                           const sx_nuts_config* cfg, const double* init,
                           const double* init_inv_metric, sx_nuts_result* out);

/* A measured HBM ceiling for roofline figures: times `iters` launches of a
 * plain copy of `bytes` (16-byte loads and stores, device to device) and
 * returns GB/s counting both the read and the write (2 * bytes per copy),
 * and the milliseconds per copy.  No counterpart in the reference.       */
int sx_device_copy_bandwidth(int device, int64_t bytes, int iters, double* gbs,
                             double* ms_per_copy);

/* ---------------- batched device-resident sampler ---------------- */
/* `chains` chains of adaptive diag_e NUTS on the device, chain ids
 * cfg->chain, cfg->chain + 1, ... (the reference's multi-chain service,
 * src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:328-407, runs one
 * thread per chain instead).  Every chain draws from its own Philox stream
 * {seed, chain id}, the stream of the single-chain service.
 * grad_precision: SX_FP64 (the batched FP64 pass) or SX_BF16X2 (MFMA;
 * needs a context created with that precision).  inits: chains x D or
 * NULL; inv_metric: D or NULL.                                          */
typedef struct sx_nuts_batch sx_nuts_batch;
int sx_nuts_batch_create(sx_nuts_batch** out, sx_glm* m,
                         const sx_nuts_config* cfg, int64_t chains,
                         int grad_precision, const double* inits,
                         const double* inv_metric);
/* Advance all chains by up to max_steps lockstep gradient steps (stops
 * early once every chain has finished).                                 */
/* Steps run as replays of a captured HIP graph of 16 steps; with the
 * environment variable STANX_BATCH_NO_GRAPH=1 at create they are launched
 * directly instead (same kernels and results; for kernel tracers).     */
int sx_nuts_batch_run(sx_nuts_batch* b, int64_t max_steps,
                      int64_t* steps_run, int64_t* chains_done);
/* Rows of draws each chain can hold ((save_warmup ? warmup : 0) + samples,
 * thinned).                                                             */
int64_t sx_nuts_batch_rows_cap(const sx_nuts_batch* b);
/* draws: chains x rows_cap x (7 + D), columns as in sx_nuts_result with
 * constrained parameters; rows, status (0 / 70 / 78, -1 still running),
 * stepsize, n_grad: per chain; inv_metric: chains x D.  Any output may be
 * NULL.                                                                 */
int sx_nuts_batch_results(sx_nuts_batch* b, double* draws, int64_t* rows,
                          int* status, double* stepsize, double* inv_metric,
                          int64_t* n_grad, double* seconds, int64_t* steps);
void sx_nuts_batch_destroy(sx_nuts_batch* b);

/* ---------------- diagnostics (host only) ---------------- */
/* x: m chains of n draws, column-major (chain c at x + c * n).
 * kind 0: stan::analyze::ess (src/stan/analyze/mcmc/ess.hpp:28-103);
 * kind 1: the bulk ESS of split_rank_normalized_ess
 * (src/stan/analyze/mcmc/split_rank_normalized_ess.hpp:28-55), NaN when
 * the split chains do not vary.  Used by the serving loop's stopping rule. */
int sx_analyze_ess(const double* x, int64_t n, int64_t m, int kind, double* out);

/* ---------------- serving ---------------- */
/* A request loop over posterior jobs on one GPU (no counterpart in the
 * reference, whose services run one request at a time).  Jobs are admitted
 * in arrival order as soon as their HBM footprint fits the budget (first
 * fit), at most max_active at once; each resident job keeps its chain state
 * on the device and its graph launches overlap with the other jobs' on its
 * own stream.  When a job's chains finish, its min bulk ESS is computed; if
 * it is below target_min_ess the sampling phase is continued (up to
 * max_samples iterations per chain), else the job completes and its device
 * memory is freed.  A job's draws do not depend on the other jobs.      */
typedef struct sx_server sx_server; /* opaque */

#define SX_JOB_QUEUED 0
#define SX_JOB_RUNNING 1
#define SX_JOB_DONE 2
#define SX_JOB_FAILED 3

typedef struct sx_job_spec {
  int family;                 /* SX_LOGIT, SX_NORMAL, SX_HIER_LOGIT            */
  const double* X;            /* N x K row-major (copied at submit)            */
  const void* y;              /* as sx_glm_create (copied)                     */
  int64_t N, K;
  const int64_t* group_offsets; /* J + 1 (SX_HIER_LOGIT), else NULL (copied) */
  int64_t J;
  sx_prior prior;
  int64_t chains;
  int grad_precision;         /* SX_BF16X2 (used when K <= 512 and chains x K
                                 exceeds the MFMA tile's cost; FP64
                                 otherwise: sx_job_info.grad_precision tells),
                                 SX_FP32 (SX_HIER_LOGIT) or SX_FP64          */
  sx_nuts_config cfg;         /* chain ids cfg.chain, cfg.chain + 1, ...;
                                 cfg.num_samples: first sampling phase       */
  double target_min_ess;      /* <= 0: run cfg.num_samples only              */
  int64_t max_samples;        /* cap on sampling iterations when extending   */
} sx_job_spec;

typedef struct sx_job_info {
  int state;                  /* SX_JOB_*                                      */
  int rc;                     /* status of a failed job                        */
  int64_t footprint_bytes;    /* HBM reserved while resident                   */
  double t_submit, t_admit, t_done; /* seconds since the server started      */
  int64_t n_grad;             /* gradient evaluations over all chains          */
  int64_t steps;              /* lockstep steps                                */
  int64_t D, chains, rows_cap;
  int64_t sampling_iterations;
  int extensions;
  double min_ess;             /* min bulk ESS over parameters at completion    */
  int grad_precision;         /* SX_FP64 / SX_FP32 / SX_BF16X2 the job runs with:
                                 BF16 hi+lo when the job allows it and it costs
                                 less than the FP64 passes                      */
} sx_job_info;

int sx_server_create(sx_server** out, int device, int64_t hbm_budget_bytes,
                     int max_active);
/* estimated device bytes of a job (what admission reserves)            */
int sx_server_footprint(const sx_job_spec* spec, int64_t* bytes);
/* non-blocking; SX_ENOMEM if the footprint exceeds the whole budget     */
int sx_server_submit(sx_server* srv, const sx_job_spec* spec, int64_t* job_id);
int sx_server_info(sx_server* srv, int64_t job_id, sx_job_info* info);
/* waits for completion (timeout_s < 0: no limit); SX_ETIMEOUT on timeout,
 * the job's status if it failed                                          */
int sx_server_wait(sx_server* srv, int64_t job_id, double timeout_s);
/* draws: chains x rows_cap x (7 + D) as sx_nuts_batch_results; rows and
 * status: per chain                                                     */
int sx_server_results(sx_server* srv, int64_t job_id, double* draws,
                      int64_t* rows, int* status);
/* forget a finished job and its host copy of the draws                  */
int sx_server_release(sx_server* srv, int64_t job_id);
/* drops queued jobs and cancels resident ones (both end SX_JOB_FAILED),
 * then stops the scheduler                                              */
void sx_server_destroy(sx_server* srv);

#ifdef __cplusplus
}
#endif
#endif
