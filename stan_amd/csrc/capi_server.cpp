// Serving: a request loop over posterior jobs on one GPU.
//
// Callers submit jobs (data, chain count, sampler configuration, a target
// min-ESS) from any thread.  One scheduler thread owns every HIP call:
//   admission   jobs leave the queue in arrival order as soon as their
//               estimated HBM footprint fits the free part of the budget
//               (first fit: a smaller later job may pass a large one that
//               does not fit yet), up to max_active resident jobs;
//   residency   an admitted job gets its own GLM context and device-resident
//               batched sampler (chain state, trees, adaptation, draws stay on
//               the device) on its own stream;
//   progress    every resident job keeps up to two graph launches (16
//               lockstep NUTS steps each) queued on its stream, so the work
//               of all resident jobs overlaps on the device;
//   completion  when every chain of a job has finished its iterations, its
//               draws come back and the min bulk ESS is computed on a host
//               thread; below the target the sampling phase is extended
//               (the same run continued, not restarted), otherwise the job's
//               device memory is freed and waiting jobs are admitted.
// A job's result depends only on its own data, seed and configuration: its
// kernels use launch shapes fixed by its sizes (no autotuning here), its
// chains draw from their own Philox streams, and the stopping rule reads
// its own draws.  So a job gives the same draws alone and in any mix.
// No counterpart in the reference, whose services run one request at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../include/stanx/analyze.hpp"
#include "nuts_batch.hpp"
#include "sx_internal.hpp"

using sx::set_error;

namespace {

using clock_t_ = std::chrono::steady_clock;

struct job {
  int64_t id = 0;
  sx_job_spec spec{};
  std::vector<double> X, yd;
  std::vector<int32_t> yi;
  std::vector<int64_t> offsets;
  int64_t footprint = 0;
  int state = SX_JOB_QUEUED;
  int rc = SX_OK;
  std::string error;
  double t_submit = 0, t_admit = 0, t_done = 0;
  sx_glm* m = nullptr;
  sx_nuts_batch* b = nullptr;
  int64_t D = 0, rows_cap = 0, samples = 0;
  int extensions = 0;
  double min_ess = NAN;
  int64_t n_grad = 0, steps = 0;
  std::vector<double> draws;
  std::vector<int64_t> rows;
  std::vector<int> status;
  std::future<double> ess;  // pending analysis
  bool analysing = false;
};

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// A job that allows BF16 hi+lo gradients gets them when they cost less
// than the exact FP64 passes.  The MFMA pass computes whole tiles of 128
// chains at about 0.09 ns + 2.5 ps x Kp per row and tile (measured at
// N = 1e6, K = 100 and N = 2e5, K = 256), and a batch of at most 64 chains
// one 64-chain tile at about 0.6 of that (0.257 vs 0.433 ms at N = 1e6,
// K = 100); above K = 256 the two-launch path costs up to about 1.5 times
// that (0.46 ms at N = 2e5, K = 320, 256 chains; 0.88 ms at N = 1e6,
// K = 300, 64 chains).  The FP64 passes stream 8 K bytes per row and
// chain, 1.25 ps x K per row and chain at 6.4 TB/s.  So the MFMA pass wins
// when chains x K > (72 + 2 Kp) per 128-chain tile, or 0.6 (72 + 2 Kp) for
// one 64-chain tile, times 1.5 above K = 256.
int gradient_precision(const sx_job_spec& s) {
  if (s.family == SX_HIER_LOGIT) return s.grad_precision == SX_FP32 ? SX_FP32 : SX_FP64;
  if (s.grad_precision != SX_BF16X2 || s.K > 512) return SX_FP64;
  const int64_t Kp = s.K <= 256 ? std::max<int64_t>(32, ceil_div(s.K, 32) * 32) : ceil_div(s.K, 64) * 64;
  const int64_t ck = s.chains * s.K, tile = 72 + 2 * Kp, wide = Kp > 256 ? 3 : 2;
  const bool mfma = s.chains <= 64 ? 10 * ck > 3 * tile * wide
                                   : 2 * ck > ceil_div(s.chains, 128) * tile * wide;
  return mfma ? SX_BF16X2 : SX_FP64;
}

// Device bytes a job holds while resident: data layouts, the sampler's
// chain state and draws, gradient workspaces (upper bounds), and the
// row-major staging copy of X that exists while the context is created.
int64_t footprint_of(const sx_job_spec& s) {
  const int64_t N = s.N, K = s.K, C = s.chains;
  const int prec = gradient_precision(s);
  int64_t D = s.family == SX_LOGIT ? K + 1 : s.family == SX_NORMAL ? K + 2
                                                                    : 2 * K + 2 + s.J * (K + 1);
  int64_t bytes = 0;
  bytes += 8 * N * K;                                             // staging (transient)
  bytes += s.family == SX_NORMAL ? 8 * N : N;                     // y
  if (s.family == SX_HIER_LOGIT) {
    bytes += 8 * (N * K + 64 * K) + (prec == SX_FP32 ? 4 * (N * K + 32 * K + 64) : 0);
    if (prec == SX_FP32 && K <= 31)  // the split copy of X (3 BF16 terms)
      bytes += 48 * (N + 32) * sx::hier_x6_chunks(static_cast<int>(K));
    const int64_t chunks = ceil_div(N, sx::kHierChunkMin) + s.J;  // any autotuned chunk size
    bytes += chunks * 16 + (s.J + 1) * 8;
    bytes += 8 * ((K + 3) * std::max<int64_t>(1, chunks) + (2 * K + 5) * s.J) * (C + 1);
    if (prec == SX_FP32 && K <= 31) bytes += 192 * s.J * (C + 1);  // split coefficients
  } else {
    int L = 1;
    while (L < 64 && ceil_div(K, L) > 32) L <<= 1;
    const int64_t R = 64 / L, S = ceil_div(K, L);
    bytes += ceil_div(N, R) * S * 64 * 8;                         // tiled FP64 X
    bytes += 2048 * (K + 3) * 8;                                  // context partials
    if (prec == SX_BF16X2) {
      const int64_t Np = ceil_div(N, 32) * 32;
      const int64_t Kp = K <= 256 ? std::max<int64_t>(32, ceil_div(K, 32) * 32) : ceil_div(K, 64) * 64;
      const int64_t Cp = ceil_div(C, 128) * 128;
      bytes += 2 * Np * Kp * 2 + 4 * Np;                          // X hi/lo, y per row
      bytes += Cp * Kp * 6 + Cp * 4 + Cp * Kp * 4 * 512 + 3 * Cp * 512 * 8;  // workspace
      if (Kp > 256) bytes += Np * Cp * 4;                         // residuals hi/lo
    } else {
      bytes += 2048 * (K + 3) * 8 * C;                            // per-chain partials
    }
  }
  const int thin = s.cfg.num_thin > 0 ? s.cfg.num_thin : 1;
  const int64_t samples = std::max<int64_t>(s.cfg.num_samples, s.max_samples);
  const int64_t rows = (s.cfg.save_warmup ? ceil_div(s.cfg.num_warmup, thin) : 0) +
                       ceil_div(samples, thin);
  const int64_t Dpad = (D + 3) & ~int64_t(3);
  bytes += C * sx::nb::kNumSlots * Dpad * 8;                      // chain vectors
  bytes += C * rows * (7 + D) * 8;                                // draws
  bytes += C * (2 * D + 2) * 8 + C * 1024;                        // theta, out, states
  return bytes + (16 << 20);
}

}  // namespace

struct sx_server {
  int device = 0;
  int64_t budget = 0, used = 0;
  int max_active = 16;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<job*> pending;
  std::vector<job*> active;
  std::map<int64_t, std::unique_ptr<job>> jobs;
  std::thread th;
  std::atomic<bool> stop{false};
  bool log = false;  // STANX_SERVER_LOG: scheduler events on stderr
  int64_t next_id = 1;
  clock_t_::time_point t0 = clock_t_::now();
  std::string fatal;
  double now() const { return std::chrono::duration<double>(clock_t_::now() - t0).count(); }
};

namespace {

void log_event(sx_server* srv, const job* j, const char* what) {
  if (!srv->log) return;
  std::fprintf(stderr, "[server %8.3f] job %lld %s (N=%lld K=%lld chains=%lld samples=%lld ext=%d ess=%.1f)\n",
               srv->now(), static_cast<long long>(j->id), what, static_cast<long long>(j->spec.N),
               static_cast<long long>(j->spec.K), static_cast<long long>(j->spec.chains),
               static_cast<long long>(j->samples), j->extensions, j->min_ess);
}

void finish_job(sx_server* srv, job* j, int rc, const std::string& err) {
  log_event(srv, j, rc == SX_OK ? "done" : "failed");
  if (j->b) {
    sx_nuts_batch_destroy(j->b);
    j->b = nullptr;
  }
  if (j->m) {
    sx_glm_destroy(j->m);
    j->m = nullptr;
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  j->rc = rc;
  j->error = err;
  j->state = rc == SX_OK ? SX_JOB_DONE : SX_JOB_FAILED;
  j->t_done = srv->now();
  srv->used -= j->footprint;
  srv->active.erase(std::find(srv->active.begin(), srv->active.end(), j));
  j->X.clear();
  j->X.shrink_to_fit();
  srv->cv.notify_all();
}

// context + batched sampler on the scheduler thread
int start_job(sx_server* srv, job* j) {
  const sx_job_spec& s = j->spec;
  const int prec = gradient_precision(s);
  const int ctx_prec = prec == SX_BF16X2 ? SX_BF16X2 : prec == SX_FP32 ? SX_FP32 : SX_FP64;
  sx_tuning tune{0, 0, 0, 0, 0, 0, 0};  // fixed launch shapes: results independent of load
  const void* y = s.family == SX_NORMAL ? static_cast<const void*>(j->yd.data())
                                        : static_cast<const void*>(j->yi.data());
  int rc = sx_glm_create_ex(&j->m, s.family, j->X.data(), y, s.N, s.K,
                            j->offsets.empty() ? nullptr : j->offsets.data(), s.J, ctx_prec,
                            &s.prior, srv->device, &tune);
  if (rc) return rc;
  j->X.clear();
  j->X.shrink_to_fit();
  int64_t D = 0;
  sx_glm_num_params(j->m, &D);
  j->D = D;
  j->samples = s.cfg.num_samples;
  // one workgroup per chain: jobs share the device, so a chain spread
  // over several workgroups could not count on their co-residency
  rc = sx::nb::create_batch(&j->b, j->m, &s.cfg, s.chains, prec, nullptr, nullptr,
                            std::max<int64_t>(s.cfg.num_samples, s.max_samples), 1);
  if (rc) return rc;
  j->rows_cap = sx_nuts_batch_rows_cap(j->b);
  return SX_OK;
}

// draws back to the host; min bulk ESS over parameters on a host thread
int collect(job* j) {
  const int64_t C = j->spec.chains, ncol = 7 + j->D;
  j->draws.assign(static_cast<size_t>(C * j->rows_cap * ncol), 0.0);
  j->rows.assign(static_cast<size_t>(C), 0);
  j->status.assign(static_cast<size_t>(C), 0);
  std::vector<int64_t> ng(static_cast<size_t>(C));
  double secs = 0;
  int64_t steps = 0;
  int rc = sx_nuts_batch_results(j->b, j->draws.data(), j->rows.data(), j->status.data(),
                                 nullptr, nullptr, ng.data(), &secs, &steps);
  if (rc) return rc;
  j->steps = steps;
  j->n_grad = 0;
  for (int64_t v : ng) j->n_grad += v;
  const int thin = j->spec.cfg.num_thin > 0 ? j->spec.cfg.num_thin : 1;
  const int64_t warm_rows =
      j->spec.cfg.save_warmup ? ceil_div(j->spec.cfg.num_warmup, thin) : 0;
  int64_t n = INT64_MAX;
  for (int64_t c = 0; c < C; ++c) n = std::min<int64_t>(n, j->rows[c] - warm_rows);
  const double* base = j->draws.data() + warm_rows * ncol;
  const size_t P = static_cast<size_t>(j->D), rows_cap = static_cast<size_t>(j->rows_cap);
  const size_t nn = static_cast<size_t>(std::max<int64_t>(0, n)), mm = static_cast<size_t>(C);
  j->ess = std::async(std::launch::async, [base, nn, mm, ncol, rows_cap, P]() {
    return stanx::analyze::min_bulk_ess(base, nn, mm, static_cast<size_t>(ncol),
                                        rows_cap * static_cast<size_t>(ncol), 7, P);
  });
  j->analysing = true;
  return SX_OK;
}

void scheduler(sx_server* srv) {
  if (hipSetDevice(srv->device) != hipSuccess) {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->fatal = "hipSetDevice failed";
    return;
  }
  while (true) {
    std::vector<job*> admit;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      if (srv->active.empty() && srv->pending.empty())
        srv->cv.wait_for(lk, std::chrono::milliseconds(5));
      if (srv->stop) {
        // queued jobs are dropped, resident ones cancelled (below)
        for (job* j : srv->pending) {
          j->state = SX_JOB_FAILED;
          j->rc = SX_EINVAL;
          j->error = "server stopped before the job was admitted";
        }
        srv->pending.clear();
        srv->cv.notify_all();
        if (srv->active.empty()) return;
      }
    }
    if (srv->stop) {
      std::vector<job*> act;
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        act = srv->active;
      }
      for (job* j : act) {
        if (j->analysing) j->ess.wait();
        finish_job(srv, j, SX_EINVAL, "cancelled: server stopped");
      }
      continue;
    }
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      for (auto it = srv->pending.begin(); it != srv->pending.end();) {
        job* j = *it;
        if (static_cast<int>(srv->active.size() + admit.size()) >= srv->max_active) break;
        if (srv->used + j->footprint <= srv->budget) {
          srv->used += j->footprint;
          j->state = SX_JOB_RUNNING;
          j->t_admit = srv->now();
          admit.push_back(j);
          it = srv->pending.erase(it);
        } else {
          ++it;
        }
      }
      for (job* j : admit) srv->active.push_back(j);
    }
    for (job* j : admit) {
      log_event(srv, j, "admitted");
      const int rc = start_job(srv, j);
      log_event(srv, j, "started");
      if (rc) finish_job(srv, j, rc, sx_last_error());
    }
    std::vector<job*> act;
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      act = srv->active;
    }
    bool progressed = !admit.empty();
    for (job* j : act) {
      if (!j->b) continue;
      if (j->analysing) {
        if (j->ess.wait_for(std::chrono::seconds(0)) != std::future_status::ready) continue;
        j->analysing = false;
        progressed = true;
        j->min_ess = j->ess.get();
        const sx_job_spec& s = j->spec;
        const bool below = s.target_min_ess > 0 && !(j->min_ess >= s.target_min_ess);
        const int64_t room = std::max<int64_t>(s.cfg.num_samples, s.max_samples) - j->samples;
        bool ok = true;
        for (int st : j->status) ok &= st == 0;
        if (below && room > 0 && ok) {
          const double ratio = std::isfinite(j->min_ess) && j->min_ess > 0
                                   ? s.target_min_ess / j->min_ess
                                   : 2.0;
          int64_t extra = static_cast<int64_t>(std::ceil(j->samples * (ratio - 1.0) * 1.2)) + 10;
          extra = std::max<int64_t>(10, std::min<int64_t>(extra, room));
          const int rc = sx::nb::extend(j->b, static_cast<int>(extra));
          if (rc) {
            finish_job(srv, j, rc, sx_last_error());
            continue;
          }
          j->samples += extra;
          ++j->extensions;
          log_event(srv, j, "extended");
          j->draws.clear();
          continue;
        }
        finish_job(srv, j, SX_OK, "");
        continue;
      }
      bool finished = false;
      int rc = sx::nb::poll(j->b, &finished);
      if (!rc && finished) {
        log_event(srv, j, "collect");
        rc = collect(j);
        progressed = true;
        if (rc) finish_job(srv, j, rc, sx_last_error());
        continue;
      }
      if (!rc && sx::nb::inflight(j->b) < 2) {
        rc = sx::nb::launch(j->b);
        progressed = true;
      }
      if (rc) finish_job(srv, j, rc, sx_last_error());
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

job* find_job(sx_server* srv, int64_t id) {
  auto it = srv->jobs.find(id);
  return it == srv->jobs.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int sx_analyze_ess(const double* x, int64_t n, int64_t m, int kind, double* out) {
  if (!x || !out || n < 1 || m < 1 || (kind != 0 && kind != 1)) {
    set_error("sx_analyze_ess: invalid argument");
    return SX_EINVAL;
  }
  *out = kind == 0 ? stanx::analyze::ess(x, static_cast<size_t>(n), static_cast<size_t>(m))
                   : stanx::analyze::bulk_ess(x, static_cast<size_t>(n), static_cast<size_t>(m));
  return SX_OK;
}

int sx_server_footprint(const sx_job_spec* spec, int64_t* bytes) {
  if (!spec || !bytes) {
    set_error("sx_server_footprint: NULL argument");
    return SX_EINVAL;
  }
  *bytes = footprint_of(*spec);
  return SX_OK;
}

int sx_server_create(sx_server** out, int device, int64_t hbm_budget_bytes, int max_active) {
  if (!out || hbm_budget_bytes <= 0 || max_active <= 0) {
    set_error("sx_server_create: invalid argument");
    return SX_EINVAL;
  }
  int ndev = 0;
  SX_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_error("sx_server_create: no such device");
    return SX_EINVAL;
  }
  sx_server* srv = new sx_server();
  srv->device = device;
  srv->budget = hbm_budget_bytes;
  srv->max_active = max_active;
  const char* lg = std::getenv("STANX_SERVER_LOG");
  srv->log = lg && lg[0] && lg[0] != '0';
  srv->th = std::thread(scheduler, srv);
  *out = srv;
  return SX_OK;
}

int sx_server_submit(sx_server* srv, const sx_job_spec* spec, int64_t* job_id) {
  if (!srv || !spec || !job_id) {
    set_error("sx_server_submit: NULL argument");
    return SX_EINVAL;
  }
  const sx_job_spec& s = *spec;
  if ((s.family != SX_LOGIT && s.family != SX_NORMAL && s.family != SX_HIER_LOGIT) ||
      s.N < 0 || s.K < 0 || s.chains < 1 || (s.N > 0 && (!s.X || !s.y)) ||
      (s.family == SX_HIER_LOGIT && (!s.group_offsets || s.J < 1))) {
    set_error("sx_server_submit: invalid job");
    return SX_EINVAL;
  }
  auto j = std::make_unique<job>();
  j->spec = s;
  j->X.assign(s.X, s.X + s.N * s.K);
  if (s.family == SX_NORMAL) {
    const double* y = static_cast<const double*>(s.y);
    j->yd.assign(y, y + s.N);
  } else {
    const int32_t* y = static_cast<const int32_t*>(s.y);
    j->yi.assign(y, y + s.N);
  }
  if (s.family == SX_HIER_LOGIT) j->offsets.assign(s.group_offsets, s.group_offsets + s.J + 1);
  j->spec.X = nullptr;
  j->spec.y = nullptr;
  j->spec.group_offsets = nullptr;
  j->footprint = footprint_of(s);
  std::lock_guard<std::mutex> lk(srv->mu);
  if (j->footprint > srv->budget) {
    set_error("sx_server_submit: the job's footprint (" + std::to_string(j->footprint) +
              " bytes) exceeds the server's HBM budget");
    return SX_ENOMEM;
  }
  j->id = srv->next_id++;
  j->t_submit = srv->now();
  *job_id = j->id;
  srv->pending.push_back(j.get());
  srv->jobs[j->id] = std::move(j);
  srv->cv.notify_all();
  return SX_OK;
}

int sx_server_info(sx_server* srv, int64_t job_id, sx_job_info* info) {
  if (!srv || !info) {
    set_error("sx_server_info: NULL argument");
    return SX_EINVAL;
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  job* j = find_job(srv, job_id);
  if (!j) {
    set_error("sx_server_info: unknown job");
    return SX_EINVAL;
  }
  info->state = j->state;
  info->rc = j->rc;
  info->footprint_bytes = j->footprint;
  info->t_submit = j->t_submit;
  info->t_admit = j->t_admit;
  info->t_done = j->t_done;
  info->n_grad = j->n_grad;
  info->steps = j->steps;
  info->D = j->D;
  info->chains = j->spec.chains;
  info->rows_cap = j->rows_cap;
  info->sampling_iterations = j->samples;
  info->extensions = j->extensions;
  info->min_ess = j->min_ess;
  info->grad_precision = gradient_precision(j->spec);
  return SX_OK;
}

int sx_server_wait(sx_server* srv, int64_t job_id, double timeout_s) {
  if (!srv) {
    set_error("sx_server_wait: NULL server");
    return SX_EINVAL;
  }
  std::unique_lock<std::mutex> lk(srv->mu);
  job* j = find_job(srv, job_id);
  if (!j) {
    set_error("sx_server_wait: unknown job");
    return SX_EINVAL;
  }
  const auto pred = [&] { return j->state == SX_JOB_DONE || j->state == SX_JOB_FAILED; };
  if (timeout_s < 0) {
    srv->cv.wait(lk, pred);
  } else if (!srv->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) {
    set_error("sx_server_wait: timed out");
    return SX_ETIMEOUT;
  }
  if (j->state == SX_JOB_FAILED) {
    set_error(j->error);
    return j->rc;
  }
  return SX_OK;
}

int sx_server_results(sx_server* srv, int64_t job_id, double* draws, int64_t* rows,
                      int* status) {
  if (!srv) {
    set_error("sx_server_results: NULL server");
    return SX_EINVAL;
  }
  std::lock_guard<std::mutex> lk(srv->mu);
  job* j = find_job(srv, job_id);
  if (!j || j->state != SX_JOB_DONE) {
    set_error("sx_server_results: unknown or unfinished job");
    return SX_EINVAL;
  }
  if (draws) std::memcpy(draws, j->draws.data(), j->draws.size() * sizeof(double));
  for (int64_t c = 0; c < j->spec.chains; ++c) {
    if (rows) rows[c] = j->rows[c];
    if (status) status[c] = j->status[c];
  }
  return SX_OK;
}

int sx_server_release(sx_server* srv, int64_t job_id) {
  if (!srv) return SX_EINVAL;
  std::lock_guard<std::mutex> lk(srv->mu);
  job* j = find_job(srv, job_id);
  if (!j || (j->state != SX_JOB_DONE && j->state != SX_JOB_FAILED)) {
    set_error("sx_server_release: unknown or unfinished job");
    return SX_EINVAL;
  }
  srv->jobs.erase(job_id);
  return SX_OK;
}

void sx_server_destroy(sx_server* srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
    srv->cv.notify_all();
  }
  if (srv->th.joinable()) srv->th.join();
  delete srv;
}

}  // extern "C"
