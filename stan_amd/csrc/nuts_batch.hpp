// Shared definitions of the device-resident batched NUTS sampler.
#ifndef STANX_NUTS_BATCH_HPP
#define STANX_NUTS_BATCH_HPP

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../include/stanx/mcmc/warmup_plan.hpp"

namespace sx {
namespace nb {

constexpr int kMaxDepth = 12;  // frames per chain (max_depth <= kMaxDepth)
constexpr int kNumSlots = 23 + 5 * kMaxDepth;
constexpr int kSlotMinv = 19;  // vector slot of the diagonal inverse metric

enum phase : int {
  PH_INIT_TRY = 0,
  PH_INIT_CHECK,
  PH_SS_A,
  PH_SS_B,
  PH_POST_ADAPT,
  PH_TRANS_BEGIN,
  PH_TRANS_INIT,
  PH_DOUBLING,
  PH_LEAF,
  PH_MERGE,
  PH_TRANS_END,
  PH_DONE
};

// Scalar state of one chain (copied to LDS by the step kernel).
struct chain_state {
  int phase, pending, status, iter, rows, init_tries, adapt_flag;
  int depth, tree_depth, n_leapfrog, divergent, fwd;
  int ss_ret, ss_iter, ss_dir, last_ok;
  int mlevel, mvalid;  // PH_MERGE: next level of the ladder above a leaf, still valid
  int spec;   // the last leaf's pass also took the next leapfrog's begin half
  int znext;  // the current point is that begin half (q at theta, p at NP)
  int stage[kMaxDepth + 1];
  unsigned w_counter;  // warm-up transitions seen by the metric windows
  int w_cursor;        // next close of the window plan not yet passed
  int iter_end;  // warm-up + sampling iterations; raised to extend a run
  int xfail;     // a cross-part exchange timed out (parts > 1)
  uint64_t rng;
  uint64_t xepoch;  // cross-part exchanges done (parts > 1)
  int64_t n_grad;
  double nom_eps, eps, H0, lsw, lsw_sub, smp, accept;
  double V_z, V_fwd, V_bck, V_samp, T_samp, V_prop, T_prop, V_init, ss_H0, last_lp;
  double sa_counter, s_bar, x_bar, mu, w_n;
  double T_leaf;  // kinetic energy at the last leaf (the final half's start)
  double lsw_init[kMaxDepth + 1], lsw_final[kMaxDepth + 1];
  double V_zpf[kMaxDepth + 1], T_zpf[kMaxDepth + 1];
};

struct batch_args {
  chain_state* states;
  double* vecs;           // C x nslot x Dpad
  double* theta;          // C x D: the point each chain needs a gradient at
  const double* out;      // C x out_stride: lp, grad[D], status
  double* draws;          // C x rows_cap x ncol
  const double* inits;    // C x D or null
  unsigned* done_count;
  int64_t out_stride, rows_cap;
  int C, D, Dpad, nslot, ncol;
  uint32_t seed, chain0;
  int num_warmup, num_samples, num_thin, save_warmup, max_depth, max_init_tries;
  double init_radius, jitter, delta, gamma, kappa, t0, max_deltaH;
  stanx::mcmc::window_plan wplan;  // metric windows (the host samplers' schedule)
  // Parts per chain: with parts > 1 each chain runs on `parts` workgroups,
  // each on a slice of the D-vectors; block totals of every reduction are
  // combined across the parts through xch (C x 2 x parts x 16 tagged
  // words).  All C x parts workgroups must be resident together (the
  // launcher bounds C x parts by the CU count); a wait that exceeds
  // spin_limit fails the chain instead of hanging.
  int parts;
  int spin_limit;
  unsigned long long* xch;
  // one word per chain, nonzero once any of its parts timed out in an
  // exchange: every part stops the chain (status 70) when it sees it
  unsigned* xfail_g;
  // failure injection for tests: waits expect a tag no part writes, so
  // every exchange times out after spin_limit polls
  int xnever;
  // phase turns a chain may take per launch before it yields (large D:
  // the vector work of a deep subtree's merges would set the step time)
  int phase_budget;
};

hipError_t launch_nuts_step(const batch_args& a, hipStream_t s);

}  // namespace nb
}  // namespace sx

struct sx_glm;
struct sx_nuts_batch;
struct sx_nuts_config;

namespace sx {
namespace nb {
// create with room for up to max_samples sampling iterations per chain
// max_parts: 1 keeps every chain on one workgroup (required where other
// work shares the device, as in the serving loop); 0 lets large-D batches
// spread each chain over up to 4 workgroups.
int create_batch(sx_nuts_batch** out, sx_glm* m, const sx_nuts_config* cfg, int64_t chains,
                 int grad_precision, const double* inits, const double* inv_metric,
                 int64_t max_samples, int max_parts);
int launch(sx_nuts_batch* b);                  // non-blocking graph launch
int poll(sx_nuts_batch* b, bool* finished);    // retire finished launches
int inflight(const sx_nuts_batch* b);
int extend(sx_nuts_batch* b, int extra);       // more sampling iterations
}  // namespace nb
}  // namespace sx
#endif
