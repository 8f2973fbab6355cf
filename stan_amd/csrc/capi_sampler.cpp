// C-ABI of the sampler services: runs the C++ NUTS with a diag_e or a
// dense_e metric, adaptive or not (stanx/services.hpp), on a GPU-backed GLM
// model or on a model given as a C callback, and returns the draws as a
// table.
#include <cstring>
#include <exception>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../../include/stanx.h"
#include "../include/stanx/callbacks.hpp"
#include "../include/stanx/gpu_glm_model.hpp"
#include "../include/stanx/model.hpp"
#include "../include/stanx/services.hpp"

namespace sx {
void set_error(const std::string& msg);
}

namespace {

// forwards every call to two writers
class tee_writer : public stanx::callbacks::writer {
 public:
  tee_writer(stanx::callbacks::writer& a, stanx::callbacks::writer& b) : a_(a), b_(b) {}
  void operator()(const std::vector<std::string>& n) override { a_(n); b_(n); }
  void operator()(const std::vector<double>& v) override { a_(v); b_(v); }
  void operator()() override { a_(); b_(); }
  void operator()(const std::string& m) override { a_(m); b_(m); }

 private:
  stanx::callbacks::writer& a_;
  stanx::callbacks::writer& b_;
};

// the configuration block at the top of a Stan CSV file, in CmdStan's
// layout (key = value lines, nesting by indentation)
void write_csv_config(std::ostream& o, const std::string& model_name, const sx_nuts_config* c,
                      bool adapt, bool dense, int sig_figs) {
  o << "# stan_version_major = 2\n# stan_version_minor = 35\n# stan_version_patch = 0\n";
  o << "# model = " << model_name << "\n";
  o << "# method = sample (Default)\n#   sample\n";
  o << "#     num_samples = " << c->num_samples << "\n";
  o << "#     num_warmup = " << c->num_warmup << "\n";
  o << "#     save_warmup = " << (c->save_warmup ? "true" : "false") << "\n";
  o << "#     thin = " << (c->num_thin > 0 ? c->num_thin : 1) << "\n";
  o << "#     adapt\n#       engaged = " << (adapt ? "true" : "false") << "\n";
  o << "#       gamma = " << c->gamma << "\n#       delta = " << c->delta << "\n";
  o << "#       kappa = " << c->kappa << "\n#       t0 = " << c->t0 << "\n";
  o << "#       init_buffer = " << c->init_buffer << "\n#       term_buffer = "
    << c->term_buffer << "\n#       window = " << c->window << "\n";
  o << "#     algorithm = hmc (Default)\n#       hmc\n#         engine = nuts (Default)\n";
  o << "#           nuts\n#             max_depth = " << c->max_depth << "\n";
  o << (dense ? "#         metric = dense_e\n" : "#         metric = diag_e (Default)\n");
  o << "#         stepsize = " << c->stepsize << "\n";
  o << "#         stepsize_jitter = " << c->stepsize_jitter << "\n";
  o << "#     num_chains = 1 (Default)\n";
  o << "# id = " << c->chain << "\n# init = " << c->init_radius << "\n";
  o << "# random\n#   seed = " << c->random_seed << "\n";
  o << "# output\n#   sig_figs = " << sig_figs << "\n";
}

template <class Model>
int run_service(const Model& model, const sx_nuts_config* cfg, const double* init,
                const double* init_inv_metric, sx_nuts_result* out, bool adapt = true,
                const char* csv_path = nullptr, int sig_figs = -1, bool dense = false,
                const char* diagnostic_path = nullptr, const char* metric_path = nullptr) {
  const char* name = dense ? (adapt ? "sx_hmc_nuts_dense_e_adapt" : "sx_hmc_nuts_dense_e")
                           : (adapt ? "sx_hmc_nuts_diag_e_adapt" : "sx_hmc_nuts_diag_e");
  if (!cfg || !out || !out->draws) {
    sx::set_error(std::string(name) + ": NULL argument");
    return SX_CODE_CONFIG;
  }
  const size_t D = model.num_params_r();
  const size_t M = dense ? D * D : D;  // metric size
  std::vector<double> init_v, minv_v;
  if (init) init_v.assign(init, init + D);
  if (init_inv_metric) minv_v.assign(init_inv_metric, init_inv_metric + M);
  stanx::callbacks::interrupt interrupt;
  stanx::callbacks::counting_logger logger;
  stanx::callbacks::writer init_writer;
  stanx::callbacks::table_writer sample_writer;
  std::ofstream csv, diag, metric;
  auto open = [&](std::ofstream& f, const char* path) {
    f.open(path);
    if (!f) {
      sx::set_error(std::string("cannot open ") + path);
      return false;
    }
    if (sig_figs > 0) f << std::setprecision(sig_figs);
    return true;
  };
  if (csv_path) {
    if (!open(csv, csv_path)) return SX_CODE_CONFIG;
    write_csv_config(csv, model.model_name(), cfg, adapt, dense, sig_figs);
  }
  if (diagnostic_path) {
    if (!open(diag, diagnostic_path)) return SX_CODE_CONFIG;
    write_csv_config(diag, model.model_name(), cfg, adapt, dense, sig_figs);
  }
  if (metric_path) {
    metric.open(metric_path);
    if (!metric) {
      sx::set_error(std::string("cannot open ") + metric_path);
      return SX_CODE_CONFIG;
    }
  }
  stanx::callbacks::stream_writer csv_writer(csv, "# "), diag_file_writer(diag, "# ");
  stanx::callbacks::writer null_writer;
  tee_writer out_writer(sample_writer, csv_path ? static_cast<stanx::callbacks::writer&>(csv_writer)
                                                : null_writer);
  stanx::callbacks::writer& diag_writer =
      diagnostic_path ? static_cast<stanx::callbacks::writer&>(diag_file_writer) : null_writer;
  stanx::callbacks::json_writer metric_json(metric);
  stanx::callbacks::structured_writer null_struct;
  stanx::callbacks::structured_writer& metric_writer =
      metric_path ? static_cast<stanx::callbacks::structured_writer&>(metric_json) : null_struct;
  double eps = 0.0, warm = 0.0, samp = 0.0;
  std::vector<double> minv_out;
  long long n_grad = 0;
  int rc = SX_CODE_SOFTWARE;
  const int thin = cfg->num_thin > 0 ? cfg->num_thin : 1;
  const std::vector<double>* iv = init ? &init_v : nullptr;
  const std::vector<double>* mv = init_inv_metric ? &minv_v : nullptr;
  try {
    namespace smp = stanx::services::sample;
    if (adapt && !dense) {
      rc = smp::hmc_nuts_diag_e_adapt(
          model, iv, mv, cfg->random_seed, cfg->chain, cfg->init_radius, cfg->num_warmup,
          cfg->num_samples, thin, cfg->save_warmup != 0, cfg->refresh, cfg->stepsize,
          cfg->stepsize_jitter, cfg->max_depth, cfg->delta, cfg->gamma, cfg->kappa, cfg->t0,
          cfg->init_buffer, cfg->term_buffer, cfg->window, interrupt, logger, init_writer,
          out_writer, diag_writer, metric_writer, &eps, &minv_out, &n_grad, &warm, &samp);
    } else if (adapt) {
      rc = smp::hmc_nuts_dense_e_adapt(
          model, iv, mv, cfg->random_seed, cfg->chain, cfg->init_radius, cfg->num_warmup,
          cfg->num_samples, thin, cfg->save_warmup != 0, cfg->refresh, cfg->stepsize,
          cfg->stepsize_jitter, cfg->max_depth, cfg->delta, cfg->gamma, cfg->kappa, cfg->t0,
          cfg->init_buffer, cfg->term_buffer, cfg->window, interrupt, logger, init_writer,
          out_writer, diag_writer, metric_writer, &eps, &minv_out, &n_grad, &warm, &samp);
    } else {
      rc = dense ? smp::hmc_nuts_dense_e(model, iv, mv, cfg->random_seed, cfg->chain,
                                         cfg->init_radius, cfg->num_warmup, cfg->num_samples,
                                         thin, cfg->save_warmup != 0, cfg->refresh,
                                         cfg->stepsize, cfg->stepsize_jitter, cfg->max_depth,
                                         interrupt, logger, init_writer, out_writer,
                                         diag_writer, &n_grad, &warm, &samp)
                 : smp::hmc_nuts_diag_e(model, iv, mv, cfg->random_seed, cfg->chain,
                                        cfg->init_radius, cfg->num_warmup, cfg->num_samples,
                                        thin, cfg->save_warmup != 0, cfg->refresh,
                                        cfg->stepsize, cfg->stepsize_jitter, cfg->max_depth,
                                        interrupt, logger, init_writer, out_writer,
                                        diag_writer, &n_grad, &warm, &samp);
      eps = cfg->stepsize;
      if (init_inv_metric) {
        minv_out = minv_v;
      } else {
        minv_out.assign(M, 0.0);
        for (size_t i = 0; i < D; ++i) minv_out[dense ? i * D + i : i] = 1.0;
      }
    }
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
  if (!logger.errors.empty()) sx::set_error(logger.errors.back());
  // rows carry: lp__, accept_stat__, 5 sampler params, then constrained
  // params from write_array
  out->rows = static_cast<int64_t>(sample_writer.rows.size());
  const size_t ncol = 7 + D;
  for (size_t i = 0; i < sample_writer.rows.size(); ++i) {
    const auto& r = sample_writer.rows[i];
    std::memcpy(out->draws + i * ncol, r.data(), sizeof(double) * std::min(r.size(), ncol));
  }
  out->stepsize = eps;
  if (out->inv_metric && minv_out.size() == M)
    std::memcpy(out->inv_metric, minv_out.data(), sizeof(double) * M);
  out->n_grad = n_grad;
  out->warmup_seconds = warm;
  out->sampling_seconds = samp;
  return rc;
}

}  // namespace

extern "C" {

int sx_hmc_nuts_diag_e_adapt(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                             const double* init_inv_metric, sx_nuts_result* out) {
  if (!m) {
    sx::set_error("sx_hmc_nuts_diag_e_adapt: NULL model");
    return SX_CODE_CONFIG;
  }
  try {
    stanx::model::gpu_glm_model model(m);
    return run_service(model, cfg, init, init_inv_metric, out);
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
}

int sx_hmc_nuts_diag_e_adapt_cb(sx_lpgrad_fn f, void* ctx, int D, const sx_nuts_config* cfg,
                                const double* init, const double* init_inv_metric,
                                sx_nuts_result* out) {
  if (!f || D < 0) {
    sx::set_error("sx_hmc_nuts_diag_e_adapt_cb: invalid model");
    return SX_CODE_CONFIG;
  }
  stanx::model::callback_model model(f, ctx, D);
  return run_service(model, cfg, init, init_inv_metric, out);
}

int sx_hmc_nuts_diag_e_adapt_csv(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                                 const double* init_inv_metric, const char* csv_path,
                                 int sig_figs, sx_nuts_result* out) {
  if (!m || !csv_path) {
    sx::set_error("sx_hmc_nuts_diag_e_adapt_csv: NULL model or path");
    return SX_CODE_CONFIG;
  }
  try {
    stanx::model::gpu_glm_model model(m);
    return run_service(model, cfg, init, init_inv_metric, out, true, csv_path, sig_figs);
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
}

int sx_hmc_nuts_diag_e_adapt_cb_csv(sx_lpgrad_fn f, void* ctx, int D, const sx_nuts_config* cfg,
                                    const double* init, const double* init_inv_metric,
                                    const char* csv_path, int sig_figs, sx_nuts_result* out) {
  if (!f || D < 0 || !csv_path) {
    sx::set_error("sx_hmc_nuts_diag_e_adapt_cb_csv: invalid model or path");
    return SX_CODE_CONFIG;
  }
  stanx::model::callback_model model(f, ctx, D);
  return run_service(model, cfg, init, init_inv_metric, out, true, csv_path, sig_figs);
}

int sx_hmc_nuts_diag_e(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                       const double* init_inv_metric, sx_nuts_result* out) {
  if (!m) {
    sx::set_error("sx_hmc_nuts_diag_e: NULL model");
    return SX_CODE_CONFIG;
  }
  try {
    stanx::model::gpu_glm_model model(m);
    return run_service(model, cfg, init, init_inv_metric, out, false);
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
}

int sx_hmc_nuts_diag_e_cb(sx_lpgrad_fn f, void* ctx, int D, const sx_nuts_config* cfg,
                          const double* init, const double* init_inv_metric,
                          sx_nuts_result* out) {
  if (!f || D < 0) {
    sx::set_error("sx_hmc_nuts_diag_e_cb: invalid model");
    return SX_CODE_CONFIG;
  }
  stanx::model::callback_model model(f, ctx, D);
  return run_service(model, cfg, init, init_inv_metric, out, false);
}

// ---- file outputs: sample CSV, diagnostic CSV, metric JSON ----
int sx_hmc_nuts_diag_e_adapt_files(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                                   const double* init_inv_metric, const sx_output_files* files,
                                   sx_nuts_result* out) {
  if (!m || !files) {
    sx::set_error("sx_hmc_nuts_diag_e_adapt_files: NULL model or files");
    return SX_CODE_CONFIG;
  }
  try {
    stanx::model::gpu_glm_model model(m);
    return run_service(model, cfg, init, init_inv_metric, out, true, files->sample_csv,
                       files->sig_figs, false, files->diagnostic_csv, files->metric_json);
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
}

int sx_hmc_nuts_diag_e_adapt_cb_files(sx_lpgrad_fn f, void* ctx, int D,
                                      const sx_nuts_config* cfg, const double* init,
                                      const double* init_inv_metric,
                                      const sx_output_files* files, sx_nuts_result* out) {
  if (!f || D < 0 || !files) {
    sx::set_error("sx_hmc_nuts_diag_e_adapt_cb_files: invalid model or files");
    return SX_CODE_CONFIG;
  }
  stanx::model::callback_model model(f, ctx, D);
  return run_service(model, cfg, init, init_inv_metric, out, true, files->sample_csv,
                     files->sig_figs, false, files->diagnostic_csv, files->metric_json);
}

int sx_hmc_nuts_diag_e_files(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                             const double* init_inv_metric, const sx_output_files* files,
                             sx_nuts_result* out) {
  if (!m || !files) {
    sx::set_error("sx_hmc_nuts_diag_e_files: NULL model or files");
    return SX_CODE_CONFIG;
  }
  try {
    stanx::model::gpu_glm_model model(m);
    return run_service(model, cfg, init, init_inv_metric, out, false, files->sample_csv,
                       files->sig_figs, false, files->diagnostic_csv, nullptr);
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
}

// ---- dense_e metric ----
int sx_hmc_nuts_dense_e_adapt(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                              const double* init_inv_metric, sx_nuts_result* out) {
  if (!m) {
    sx::set_error("sx_hmc_nuts_dense_e_adapt: NULL model");
    return SX_CODE_CONFIG;
  }
  try {
    stanx::model::gpu_glm_model model(m);
    return run_service(model, cfg, init, init_inv_metric, out, true, nullptr, -1, true);
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
}

int sx_hmc_nuts_dense_e_adapt_cb(sx_lpgrad_fn f, void* ctx, int D, const sx_nuts_config* cfg,
                                 const double* init, const double* init_inv_metric,
                                 sx_nuts_result* out) {
  if (!f || D < 0) {
    sx::set_error("sx_hmc_nuts_dense_e_adapt_cb: invalid model");
    return SX_CODE_CONFIG;
  }
  stanx::model::callback_model model(f, ctx, D);
  return run_service(model, cfg, init, init_inv_metric, out, true, nullptr, -1, true);
}

int sx_hmc_nuts_dense_e_adapt_csv(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                                  const double* init_inv_metric, const char* csv_path,
                                  int sig_figs, sx_nuts_result* out) {
  if (!m || !csv_path) {
    sx::set_error("sx_hmc_nuts_dense_e_adapt_csv: NULL model or path");
    return SX_CODE_CONFIG;
  }
  try {
    stanx::model::gpu_glm_model model(m);
    return run_service(model, cfg, init, init_inv_metric, out, true, csv_path, sig_figs, true);
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
}

int sx_hmc_nuts_dense_e_adapt_cb_csv(sx_lpgrad_fn f, void* ctx, int D, const sx_nuts_config* cfg,
                                     const double* init, const double* init_inv_metric,
                                     const char* csv_path, int sig_figs, sx_nuts_result* out) {
  if (!f || D < 0 || !csv_path) {
    sx::set_error("sx_hmc_nuts_dense_e_adapt_cb_csv: invalid model or path");
    return SX_CODE_CONFIG;
  }
  stanx::model::callback_model model(f, ctx, D);
  return run_service(model, cfg, init, init_inv_metric, out, true, csv_path, sig_figs, true);
}

int sx_hmc_nuts_dense_e(sx_glm* m, const sx_nuts_config* cfg, const double* init,
                        const double* init_inv_metric, sx_nuts_result* out) {
  if (!m) {
    sx::set_error("sx_hmc_nuts_dense_e: NULL model");
    return SX_CODE_CONFIG;
  }
  try {
    stanx::model::gpu_glm_model model(m);
    return run_service(model, cfg, init, init_inv_metric, out, false, nullptr, -1, true);
  } catch (const std::exception& e) {
    sx::set_error(e.what());
    return SX_CODE_SOFTWARE;
  }
}

int sx_hmc_nuts_dense_e_cb(sx_lpgrad_fn f, void* ctx, int D, const sx_nuts_config* cfg,
                           const double* init, const double* init_inv_metric,
                           sx_nuts_result* out) {
  if (!f || D < 0) {
    sx::set_error("sx_hmc_nuts_dense_e_cb: invalid model");
    return SX_CODE_CONFIG;
  }
  stanx::model::callback_model model(f, ctx, D);
  return run_service(model, cfg, init, init_inv_metric, out, false, nullptr, -1, true);
}

}  // extern "C"
