// C-ABI of the device-resident batched sampler: many chains of adaptive
// diag_e NUTS on one GLM context, all state on the device.  A run is a
// sequence of lockstep steps (nuts_step, then one batched gradient), and
// fixed runs of steps are captured once in a HIP graph and replayed until
// every chain has finished.  This replaces the reference's one-thread-per-
// chain loop (src/stan/services/sample/hmc_nuts_diag_e_adapt.hpp:328-407).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nuts_batch.hpp"
#include "sx_internal.hpp"

using sx::set_error;

struct sx_nuts_batch {
  sx_glm* m = nullptr;
  int64_t C = 0, D = 0, Dpad = 0, rows_cap = 0;
  int grad_precision = SX_FP64;
  int steps_per_graph = 16;
  hipStream_t stream = nullptr;
  sx::nb::batch_args a{};
  sx::pass_args g{};     // batched gradient arguments (FP64 path)
  int grad_blocks = 1;   // blocks per chain of the FP64 batched pass
  double* d_partials = nullptr;
  double* d_out = nullptr;
  double* d_theta = nullptr;
  double* d_inits = nullptr;
  sx::mx_workspace mx;     // this sampler's MFMA workspace
  sx::hier_workspace hw;   // this sampler's hierarchical partial sums
  unsigned* h_done = nullptr;  // pinned
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t steps = 0;
  double seconds = 0.0;
  bool finished = false;
  // non-blocking driving (serving): graph launches in flight, each followed
  // by a copy of the done counter and an event
  static constexpr int kRing = 2;
  hipEvent_t ev[kRing] = {nullptr, nullptr};
  int head = 0, inflight = 0;
};

namespace {

int enqueue_gradient(sx_nuts_batch* b) {
  if (b->m->family == SX_HIER_LOGIT)
    return sx::hier_enqueue(b->m, b->hw, b->d_theta, static_cast<int>(b->C), b->d_out, 1, 1,
                            b->grad_precision == SX_FP32, b->stream);
  if (b->grad_precision == SX_BF16X2)
    return sx::mx_enqueue(b->m, b->mx, b->d_theta, static_cast<int>(b->C), b->d_out, 1, 1,
                          b->stream);
  SX_HIP(sx::launch_pass_f64(b->m->family, b->m->geom, b->g, b->grad_blocks, b->m->threads,
                             b->stream, true));
  return SX_OK;
}

int enqueue_step(sx_nuts_batch* b) {
  SX_HIP(sx::nb::launch_nuts_step(b->a, b->stream));
  return enqueue_gradient(b);
}

void release(sx_nuts_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipFree(b->a.states);
  (void)hipFree(b->a.vecs);
  (void)hipFree(b->a.draws);
  (void)hipFree(b->a.done_count);
  (void)hipFree(b->a.xch);
  (void)hipFree(b->a.xfail_g);
  (void)hipFree(b->d_partials);
  (void)hipFree(b->d_out);
  (void)hipFree(b->d_theta);
  (void)hipFree(b->d_inits);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  sx::free_workspace(b->mx);
  sx::free_workspace(b->hw);
  if (b->h_done) (void)hipHostFree(b->h_done);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace

namespace sx {
namespace nb {

int create_batch(sx_nuts_batch** out, sx_glm* m, const sx_nuts_config* cfg, int64_t chains,
                 int grad_precision, const double* inits, const double* inv_metric,
                 int64_t max_samples, int max_parts) {
  if (!out || !m || !cfg || chains <= 0) {
    set_error("sx_nuts_batch_create: invalid argument");
    return SX_EINVAL;
  }
  *out = nullptr;
  if (grad_precision != SX_FP64 && !(grad_precision == SX_BF16X2 && m->precision == SX_BF16X2) &&
      !(grad_precision == SX_FP32 && m->precision == SX_FP32)) {
    set_error("sx_nuts_batch_create: mixed-precision gradients need a context created with the "
              "same precision");
    return SX_EINVAL;
  }
  const int max_depth = cfg->max_depth > 0 ? cfg->max_depth : 5;
  if (max_depth > sx::nb::kMaxDepth) {
    set_error("sx_nuts_batch_create: max_depth above " + std::to_string(sx::nb::kMaxDepth));
    return SX_EINVAL;
  }
  const int64_t D = m->D;
  if (inv_metric)
    for (int64_t i = 0; i < D; ++i)
      if (!(inv_metric[i] > 0) || !std::isfinite(inv_metric[i])) {
        set_error("Inverse metric must be positive and finite");
        return SX_EINVAL;
      }
  SX_HIP(hipSetDevice(m->device));
  sx_nuts_batch* b = new sx_nuts_batch();
  b->m = m;
  b->C = chains;
  b->D = D;
  b->Dpad = (D + 3) & ~int64_t(3);
  b->grad_precision = grad_precision;
  const int thin = cfg->num_thin > 0 ? cfg->num_thin : 1;
  const int64_t warm_rows = cfg->save_warmup ? (cfg->num_warmup + thin - 1) / thin : 0;
  const int64_t cap_samples = std::max<int64_t>(std::max(0, cfg->num_samples), max_samples);
  b->rows_cap = warm_rows + (cap_samples + thin - 1) / thin;
  auto fail = [&](int code, const char* what) {
    set_error(std::string("sx_nuts_batch_create: ") + what);
    release(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(SX_EDEVICE, "stream");
  for (auto& e : b->ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return fail(SX_EDEVICE, "event");
  const int ncol = static_cast<int>(7 + D);
  const size_t nvec = static_cast<size_t>(chains) * sx::nb::kNumSlots * b->Dpad;
  if (hipMalloc(&b->a.states, sizeof(sx::nb::chain_state) * chains) != hipSuccess ||
      hipMalloc(&b->a.vecs, sizeof(double) * nvec) != hipSuccess ||
      hipMalloc(&b->a.draws, sizeof(double) * std::max<int64_t>(1, chains * b->rows_cap) * ncol) !=
          hipSuccess ||
      hipMalloc(&b->a.done_count, sizeof(unsigned)) != hipSuccess ||
      hipMalloc(&b->d_theta, sizeof(double) * chains * D) != hipSuccess ||
      hipMalloc(&b->d_out, sizeof(double) * chains * (D + 2)) != hipSuccess ||
      hipHostMalloc(&b->h_done, sizeof(unsigned)) != hipSuccess)
    return fail(SX_ENOMEM, "device memory");
  // Parts per chain: a chain with D > 8192 (one 512-thread workgroup per
  // chain streams its vectors through one CU) may spread over up to 4
  // workgroups, as long as every workgroup of the launch can be resident
  // at once: chains x parts <= CUs (one 512-thread workgroup per CU).
  // STANX_NUTS_PARTS=n sets the count (1 disables).
  {
    int parts = 1;
    if (max_parts != 1 && D > 8192) {
      int cus = 0;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, m->device) !=
          hipSuccess)
        cus = 0;
      parts = static_cast<int>(std::min<int64_t>(4, cus / std::max<int64_t>(1, chains)));
      if (const char* e = std::getenv("STANX_NUTS_PARTS")) parts = std::atoi(e);
      if (max_parts > 0) parts = std::min(parts, max_parts);
      if (cus > 0 && chains * parts > cus) parts = static_cast<int>(cus / chains);
      parts = std::max(1, std::min(parts, 4));
    }
    b->a.parts = parts;
    b->a.spin_limit = 1 << 22;
    // Phase turns per chain and launch: above 8192 coordinates a chain's
    // vector passes cost about as much as its gradient, and one chain
    // closing a deep subtree (a merge pass per level, then the transition)
    // would hold the whole step; such a chain yields after 4 turns and
    // continues in the next step.  Below, a yield's extra step (a batched
    // gradient) costs more than the wait.  Measured totals (ms/step x
    // steps): hierarchical D = 21,042, 64 chains: 1008 unbounded, 954
    // with 6, 908 with 4, 950 with 3; logistic D = 101, 256 chains: 2508
    // unbounded, 2807 with 4.  STANX_NUTS_PHASE_BUDGET=n sets it (n >= 1).
    b->a.phase_budget = D > 8192 ? 4 : 4096;
    if (const char* e = std::getenv("STANX_NUTS_PHASE_BUDGET"))
      b->a.phase_budget = std::max(1, std::atoi(e));
    if (parts > 1) {
      // tag 0 (the cleared words) never matches: tags are epoch + 1
      const size_t nx = static_cast<size_t>(chains) * 2 * 16 * parts;
      if (hipMalloc(&b->a.xch, sizeof(unsigned long long) * nx) != hipSuccess)
        return fail(SX_ENOMEM, "exchange buffers");
      if (hipMemsetAsync(b->a.xch, 0, sizeof(unsigned long long) * nx, b->stream) != hipSuccess)
        return fail(SX_EDEVICE, "clear");
      if (hipMalloc(&b->a.xfail_g, sizeof(unsigned) * chains) != hipSuccess)
        return fail(SX_ENOMEM, "exchange failure words");
      if (hipMemsetAsync(b->a.xfail_g, 0, sizeof(unsigned) * chains, b->stream) != hipSuccess)
        return fail(SX_EDEVICE, "clear");
      // failure injection (tests): every exchange waits for a tag no part
      // writes and gives up after a short bounded wait, so each chain ends
      // with status 70 while every launch still runs to completion
      if (const char* e = std::getenv("STANX_NUTS_XFAIL_TEST")) {
        if (std::atoi(e) != 0) {
          b->a.xnever = 1;
          b->a.spin_limit = 256;
        }
      }
    }
  }
  if (hipMemsetAsync(b->a.vecs, 0, sizeof(double) * nvec, b->stream) != hipSuccess ||
      hipMemsetAsync(b->a.done_count, 0, sizeof(unsigned), b->stream) != hipSuccess ||
      hipMemsetAsync(b->d_theta, 0, sizeof(double) * chains * D, b->stream) != hipSuccess ||
      hipMemsetAsync(b->d_out, 0, sizeof(double) * chains * (D + 2), b->stream) != hipSuccess)
    return fail(SX_EDEVICE, "clear");
  if (inits) {
    if (hipMalloc(&b->d_inits, sizeof(double) * chains * D) != hipSuccess)
      return fail(SX_ENOMEM, "inits");
    if (hipMemcpy(b->d_inits, inits, sizeof(double) * chains * D, hipMemcpyHostToDevice) !=
        hipSuccess)
      return fail(SX_EDEVICE, "inits copy");
  }
  // metric slot of every chain
  {
    std::vector<double> minv(static_cast<size_t>(D), 1.0);
    if (inv_metric) std::copy(inv_metric, inv_metric + D, minv.begin());
    for (int64_t c = 0; c < chains; ++c) {
      double* dst = b->a.vecs + (c * sx::nb::kNumSlots + sx::nb::kSlotMinv) * b->Dpad;
      if (hipMemcpyAsync(dst, minv.data(), sizeof(double) * D, hipMemcpyHostToDevice,
                         b->stream) != hipSuccess)
        return fail(SX_EDEVICE, "metric copy");
    }
  }
  // chain scalar state
  // the metric windows: the schedule the host samplers compute
  stanx::callbacks::logger quiet;
  const stanx::mcmc::window_plan wplan = stanx::mcmc::plan_windows(
      static_cast<unsigned>(std::max(0, cfg->num_warmup)), cfg->init_buffer, cfg->term_buffer,
      cfg->window, "variance", quiet);
  {
    std::vector<sx::nb::chain_state> st(static_cast<size_t>(chains));
    for (auto& s : st) {
      std::memset(&s, 0, sizeof(s));
      s.phase = sx::nb::PH_INIT_TRY;
      s.nom_eps = cfg->stepsize > 0 ? cfg->stepsize : 0.1;
      s.eps = s.nom_eps;
      s.mu = std::log(10 * cfg->stepsize);
      s.iter_end = std::max(0, cfg->num_warmup) + std::max(0, cfg->num_samples);
    }
    if (hipMemcpy(b->a.states, st.data(), sizeof(sx::nb::chain_state) * chains,
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(SX_EDEVICE, "state copy");
  }
  sx::nb::batch_args& a = b->a;
  a.theta = b->d_theta;
  a.out = b->d_out;
  a.inits = b->d_inits;
  a.out_stride = D + 2;
  a.rows_cap = b->rows_cap;
  a.C = static_cast<int>(chains);
  a.D = static_cast<int>(D);
  a.Dpad = static_cast<int>(b->Dpad);
  a.nslot = sx::nb::kNumSlots;
  a.ncol = ncol;
  a.seed = cfg->random_seed;
  a.chain0 = cfg->chain;
  a.num_warmup = std::max(0, cfg->num_warmup);
  a.num_samples = std::max(0, cfg->num_samples);
  a.num_thin = thin;
  a.save_warmup = cfg->save_warmup;
  a.max_depth = max_depth;
  a.max_init_tries = (inits || cfg->init_radius == 0.0) ? 1 : 100;
  a.init_radius = cfg->init_radius;
  a.jitter = (cfg->stepsize_jitter > 0 && cfg->stepsize_jitter < 1) ? cfg->stepsize_jitter : 0.0;
  a.delta = (cfg->delta > 0 && cfg->delta < 1) ? cfg->delta : 0.5;
  a.gamma = cfg->gamma > 0 ? cfg->gamma : 0.05;
  a.kappa = cfg->kappa > 0 ? cfg->kappa : 0.75;
  a.t0 = cfg->t0 > 0 ? cfg->t0 : 10;
  a.max_deltaH = 1000;
  a.wplan = wplan;
  // batched FP64 gradient: every chain's pass on grid.y
  if (m->family != SX_HIER_LOGIT) {
    // each chain's pass uses the launch shape of the single-chain pass: a
    // chain's gradient is bitwise the plain path's, and independent of the
    // other chains in the batch
    b->grad_blocks = m->blocks;
    const size_t E = static_cast<size_t>(m->K) + 3;
    if (hipMalloc(&b->d_partials, sizeof(double) * E * b->grad_blocks * chains) != hipSuccess)
      return fail(SX_ENOMEM, "partials");
    sx::pass_args& g = b->g;
    g.X = m->d_X;
    g.y = m->d_y;
    g.theta = b->d_theta;
    g.partials = b->d_partials;
    g.out = b->d_out;
    g.N = m->N;
    g.T = m->geom.T;
    g.K = static_cast<int>(m->K);
    g.L = m->geom.L;
    g.lgL = m->geom.lgL;
    g.S = m->geom.S;
    g.D = static_cast<int>(D);
    g.E = static_cast<int>(E);
    g.propto = 1;
    g.jacobian = 1;
    g.chains = static_cast<int>(chains);
    g.theta_stride = D;
    g.out_stride = D + 2;
    g.alpha_sd = m->prior.alpha_sd;
    g.beta_sd = m->prior.beta_sd;
    g.sigma_sd = m->prior.sigma_sd;
    g.sigma_mu = m->prior.sigma_mu;
    const double h = 0.91893853320467274178032973640562;
    g.alpha_const = g.alpha_sd > 0 ? -std::log(g.alpha_sd) - h : 0.0;
    g.beta_const = g.beta_sd > 0 ? -std::log(g.beta_sd) - h : 0.0;
    g.sigma_const = g.sigma_sd > 0 ? -std::log(g.sigma_sd) - h : 0.0;
  }
  if (m->family == SX_HIER_LOGIT) {
    // allocate the partial sums outside the captured region
    const int rcp = sx::hier_prepare(m, b->hw, static_cast<int>(chains));
    if (rcp != SX_OK) return fail(rcp, "hierarchical workspace");
  } else if (grad_precision == SX_BF16X2) {
    // allocate the MFMA workspace outside the captured region
    const int rcp = sx::mx_enqueue(m, b->mx, b->d_theta, static_cast<int>(chains), b->d_out, 1,
                                   1, b->stream);
    if (rcp != SX_OK) return fail(rcp, "MFMA workspace");
  }
  // capture a fixed run of lockstep steps in a graph (or, with
  // STANX_BATCH_NO_GRAPH set, launch the steps directly: for kernel tracers)
  if (hipStreamSynchronize(b->stream) != hipSuccess) return fail(SX_EDEVICE, "sync");
  const char* no_graph = std::getenv("STANX_BATCH_NO_GRAPH");
  if (no_graph && no_graph[0] && no_graph[0] != '0') {
    *out = b;
    return SX_OK;
  }
  if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(SX_EDEVICE, "begin capture");
  int rc = SX_OK;
  for (int i = 0; i < b->steps_per_graph && rc == SX_OK; ++i) rc = enqueue_step(b);
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(b->stream, &graph);
  b->graph = graph;
  if (rc != SX_OK || ce != hipSuccess) return fail(SX_EDEVICE, "capture");
  if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess)
    return fail(SX_EDEVICE, "graph instantiate");
  *out = b;
  return SX_OK;
}

// One graph launch (steps_per_graph lockstep steps) without waiting, then
// the done counter copy and an event; at most kRing launches in flight.
int run_steps(sx_nuts_batch* b) {
  if (b->exec) {
    SX_HIP(hipGraphLaunch(b->exec, b->stream));
    return SX_OK;
  }
  for (int i = 0; i < b->steps_per_graph; ++i) {
    const int rc = enqueue_step(b);
    if (rc) return rc;
  }
  return SX_OK;
}

int launch(sx_nuts_batch* b) {
  // nothing left to do once every chain has finished (a run is resumed by
  // extend); a full ring waits for poll
  if (b->finished || b->inflight >= sx_nuts_batch::kRing) return SX_OK;
  {
    const int rc = run_steps(b);
    if (rc) return rc;
  }
  SX_HIP(hipMemcpyAsync(b->h_done, b->a.done_count, sizeof(unsigned), hipMemcpyDeviceToHost,
                        b->stream));
  const int slot = (b->head + b->inflight) % sx_nuts_batch::kRing;
  SX_HIP(hipEventRecord(b->ev[slot], b->stream));
  ++b->inflight;
  b->steps += b->steps_per_graph;
  return SX_OK;
}

// Retire completed launches; *finished once every chain is done.
int poll(sx_nuts_batch* b, bool* finished) {
  while (b->inflight > 0) {
    const hipError_t e = hipEventQuery(b->ev[b->head]);
    if (e == hipErrorNotReady) break;
    SX_HIP(e);
    b->head = (b->head + 1) % sx_nuts_batch::kRing;
    --b->inflight;
    if (*b->h_done >= static_cast<unsigned>(b->C)) b->finished = true;
  }
  *finished = b->finished && b->inflight == 0;
  return SX_OK;
}

int inflight(const sx_nuts_batch* b) { return b->inflight; }

// Raise the sampling iterations of every chain that finished normally by
// `extra` and resume it where it stopped: the continued run is the run
// that would have had the larger num_samples from the start.
int extend(sx_nuts_batch* b, int extra) {
  SX_HIP(hipStreamSynchronize(b->stream));
  b->inflight = 0;
  std::vector<chain_state> st(static_cast<size_t>(b->C));
  SX_HIP(hipMemcpy(st.data(), b->a.states, sizeof(chain_state) * b->C, hipMemcpyDeviceToHost));
  unsigned still_done = 0;
  for (auto& s : st) {
    if (s.phase == PH_DONE && s.status == 0) {
      s.iter_end += extra;
      s.phase = PH_TRANS_BEGIN;
      s.pending = 0;
    } else if (s.phase == PH_DONE) {
      ++still_done;
    }
  }
  SX_HIP(hipMemcpy(b->a.states, st.data(), sizeof(chain_state) * b->C, hipMemcpyHostToDevice));
  SX_HIP(hipMemcpy(b->a.done_count, &still_done, sizeof(unsigned), hipMemcpyHostToDevice));
  b->finished = still_done >= static_cast<unsigned>(b->C);
  return SX_OK;
}

int64_t sampling_rows(const sx_nuts_batch* b, int64_t* warm_rows) {
  const int thin = b->a.num_thin;
  *warm_rows = b->a.save_warmup ? (b->a.num_warmup + thin - 1) / thin : 0;
  return b->rows_cap;
}

hipStream_t stream(const sx_nuts_batch* b) { return b->stream; }

}  // namespace nb
}  // namespace sx

extern "C" {

int sx_nuts_batch_create(sx_nuts_batch** out, sx_glm* m, const sx_nuts_config* cfg,
                         int64_t chains, int grad_precision, const double* inits,
                         const double* inv_metric) {
  if (!cfg) {
    set_error("sx_nuts_batch_create: invalid argument");
    return SX_EINVAL;
  }
  return sx::nb::create_batch(out, m, cfg, chains, grad_precision, inits, inv_metric,
                              cfg->num_samples, 0);
}

void sx_nuts_batch_destroy(sx_nuts_batch* b) { release(b); }

int64_t sx_nuts_batch_rows_cap(const sx_nuts_batch* b) { return b ? b->rows_cap : 0; }

int sx_nuts_batch_run(sx_nuts_batch* b, int64_t max_steps, int64_t* steps_run,
                      int64_t* chains_done) {
  if (!b) {
    set_error("sx_nuts_batch_run: NULL batch");
    return SX_EINVAL;
  }
  SX_HIP(hipSetDevice(b->m->device));
  const auto t0 = std::chrono::steady_clock::now();
  int64_t run = 0;
  unsigned done = 0;
  while (run < max_steps && !b->finished) {
    {
      const int rc = sx::nb::run_steps(b);
      if (rc) return rc;
    }
    run += b->steps_per_graph;
    SX_HIP(hipMemcpyAsync(b->h_done, b->a.done_count, sizeof(unsigned), hipMemcpyDeviceToHost,
                          b->stream));
    SX_HIP(hipStreamSynchronize(b->stream));
    done = *b->h_done;
    if (done >= static_cast<unsigned>(b->C)) b->finished = true;
  }
  b->steps += run;
  b->seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (steps_run) *steps_run = run;
  if (chains_done) *chains_done = done;
  return SX_OK;
}

int sx_nuts_batch_results(sx_nuts_batch* b, double* draws, int64_t* rows, int* status,
                          double* stepsize, double* inv_metric, int64_t* n_grad,
                          double* seconds, int64_t* steps) {
  if (!b) {
    set_error("sx_nuts_batch_results: NULL batch");
    return SX_EINVAL;
  }
  SX_HIP(hipSetDevice(b->m->device));
  SX_HIP(hipStreamSynchronize(b->stream));
  std::vector<sx::nb::chain_state> st(static_cast<size_t>(b->C));
  SX_HIP(hipMemcpy(st.data(), b->a.states, sizeof(sx::nb::chain_state) * b->C,
                   hipMemcpyDeviceToHost));
  const int64_t ncol = 7 + b->D;
  if (draws) {
    SX_HIP(hipMemcpy(draws, b->a.draws, sizeof(double) * b->C * b->rows_cap * ncol,
                     hipMemcpyDeviceToHost));
    // write_array: constrained parameters (sigma = exp(u) for the normal
    // GLM, tau_k = exp(v_k) for the hierarchical model)
    int64_t c0 = 0, c1 = 0;
    if (b->m->family == SX_NORMAL) {
      c0 = ncol - 1;
      c1 = ncol;
    } else if (b->m->family == SX_HIER_LOGIT) {
      c0 = 7 + 1 + b->m->K;
      c1 = c0 + b->m->K + 1;
    }
    for (int64_t c = 0; c < b->C; ++c)
      for (int64_t r = 0; r < st[c].rows; ++r)
        for (int64_t j = c0; j < c1; ++j) {
          double& u = draws[(c * b->rows_cap + r) * ncol + j];
          u = std::exp(u);
        }
  }
  for (int64_t c = 0; c < b->C; ++c) {
    if (rows) rows[c] = st[c].rows;
    if (status) status[c] = st[c].phase == sx::nb::PH_DONE ? st[c].status : -1;
    if (stepsize) stepsize[c] = st[c].nom_eps;
    if (n_grad) n_grad[c] = st[c].n_grad;
  }
  if (inv_metric)
    SX_HIP(hipMemcpy2D(inv_metric, sizeof(double) * b->D,
                       b->a.vecs + sx::nb::kSlotMinv * b->Dpad, sizeof(double) * sx::nb::kNumSlots * b->Dpad,
                       sizeof(double) * b->D, b->C, hipMemcpyDeviceToHost));
  if (seconds) *seconds = b->seconds;
  if (steps) *steps = b->steps;
  return SX_OK;
}

}  // extern "C"
