// Device-resident adaptive diag_e NUTS for many chains at once.
//
// Every chain runs the same iterative state machine as the host sampler
// (stanx/mcmc/base_nuts.hpp), suspended whenever it needs a gradient: one
// launch of nuts_step advances each chain (one block per chain: one wave
// for D <= 2048, four or eight waves above; vector work is strided over the
// block's threads, scalar chain state lives in LDS and only thread 0 writes
// it, between barriers) from the gradient it just received to the point where it needs the next one,
// writes that point to theta[chain], and the batched gradient launches
// evaluate all chains together.  Chains therefore move in lockstep one
// gradient at a time but are otherwise independent: each owns its trees,
// adaptation and Philox stream (key {seed, chain id}, the host sampler's
// stream), so a chain's draws do not depend on which chains run beside it.
//
// Semantics followed (paths relative to the stan-dev/stan 2.35.0 tree):
//   initialisation        src/stan/services/util/initialize.hpp:68-227
//   step-size heuristic   src/stan/mcmc/hmc/base_hmc.hpp:93-161
//   transition            src/stan/mcmc/hmc/nuts/base_nuts.hpp:78-204
//   build_tree            src/stan/mcmc/hmc/nuts/base_nuts.hpp:247-352
//   leapfrog              src/stan/mcmc/hmc/integrators/expl_leapfrog.hpp:16-32
//   adaptation            src/stan/mcmc/hmc/nuts/adapt_diag_e_nuts.hpp:25-48,
//                         src/stan/mcmc/stepsize_adaptation.hpp:55-73,
//                         src/stan/mcmc/windowed_adaptation.hpp:80-110,
//                         src/stan/mcmc/var_adaptation.hpp:17-46
//   service loop          src/stan/services/util/run_adaptive_sampler.hpp:46-103
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../include/stanx/philox.hpp"
#include "nuts_batch.hpp"

#pragma clang fp contract(off)

namespace sx {
namespace nb {

namespace {

constexpr double kInf = __builtin_huge_val();
constexpr double kLog08 = -0.22314355131420976;  // log(0.8)

// ---- vector slots of a chain ----
enum : int {
  ZQ = 0, ZP, ZG,          // current point
  FQ, FP, FG,              // forward end of the trajectory
  BQ, BP, BG,              // backward end
  SQ, PQ,                  // sample, top-level proposal (positions)
  PFB, PBF,                // p at the backward end of the forward subtree / forward end of the backward subtree
  RHO, RHOF, RHOB,
  IQ, IP, IG,              // z_init of the step-size heuristic
  MINV, WM, WM2,           // metric and Welford accumulators
  NP,                      // momentum of the next leapfrog's begin half (leaf pass)
  LVL0                     // per level l >= 1: PIE, RI, ZPF, PFBG, RF
};
static_assert(MINV == kSlotMinv, "metric slot");
static_assert(LVL0 + 5 * kMaxDepth == kNumSlots, "slot count");
__device__ __forceinline__ int PIE(int l) { return LVL0 + 5 * (l - 1); }
__device__ __forceinline__ int RI(int l) { return LVL0 + 5 * (l - 1) + 1; }
__device__ __forceinline__ int ZPF(int l) { return LVL0 + 5 * (l - 1) + 2; }
__device__ __forceinline__ int PFBG(int l) { return LVL0 + 5 * (l - 1) + 3; }
__device__ __forceinline__ int RF(int l) { return LVL0 + 5 * (l - 1) + 4; }

// scalar slots of proposal-like points (position stored in a vector slot)
// id -1: top-level proposal PQ, 0: sample SQ, l >= 1: ZPF(l)
struct dest {
  int zprop;  // proposal id (-1 top, l >= 1 frame)
  int pbeg, pend, rho;
  int lsw;    // -1 top subtree accumulator, 2l: lsw_init[l], 2l+1: lsw_final[l]
};

struct ctx {
  const batch_args* a;
  chain_state* s;
  double* v;  // chain vector base
  int c, lane;  // lane = thread index in the chain's block
  int nt;       // threads per chain
  double* red;  // LDS scratch for block reductions (one slot per wave)
  uint32_t chain_id;
  const int* sl;  // LDS: the part's slice [sl[0], sl[1]) of the D-vectors, sl[2] = part
  __device__ int d0() const { return sl[0]; }
  __device__ int d1() const { return sl[1]; }
  __device__ int part() const { return sl[2]; }
  __device__ double* slot(int i) const { return v + static_cast<int64_t>(i) * a->Dpad; }
};

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) x += __shfl_xor(x, off, 64);
  return x;
}
// Sum of NS block totals over the chain's parts, in part order, so every
// part gets the same values.  Wave 0 publishes this part's totals and
// waits for every part's, one word per lane, so that the waits overlap
// instead of following each other.  Each total travels as two 64-bit
// words, (tag << 32) | one
// 32-bit half of the double, written and read with relaxed agent-scope
// atomics: a word that carries this exchange's tag (epoch + 1, mod 2^32)
// also carries its half (single-copy atomicity), so no release/acquire
// ordering, and with it no L2 writeback or invalidation, is needed.  The
// words of consecutive exchanges alternate between two buffers: a part
// rewrites a buffer only after every part has published the exchange in
// between, by which time all of them have read the buffer.  A wait longer
// than spin_limit marks the exchange failed (the step loop then ends the
// chain) rather than hanging.  Called in block-uniform control flow by every
// part in the same sequence.
template <int NS>
__device__ void part_sum(const ctx& k, double (&x)[NS]) {
  static_assert(NS <= 8, "at most 8 totals per exchange");
  const batch_args* a = k.a;
  const int P = a->parts;  // <= 4: 2 NS P <= 64 words, one per lane of wave 0
  chain_state* s = k.s;
  __syncthreads();
  // a part that has already failed an exchange neither publishes nor waits
  // again: its sums are void and the chain ends with this phase turn
  if (k.lane < 64 && !s->xfail) {
    const int ln = k.lane;
    const uint64_t ep = s->xepoch;
    const unsigned long long tag = ((ep + 1) & 0xffffffffull) << 32;
    const unsigned long long want = a->xnever ? (tag ^ (1ull << 63)) : tag;
    unsigned long long* base = a->xch + (static_cast<int64_t>(k.c) * 2 + (ep & 1)) * 16 * P;
    // publish: lane q < 2 NS writes word q of this part (half q % 2 of total
    // q / 2); the totals reach the lanes through LDS (every thread holds
    // them, lane 0 stores them with fixed indices)
    if (ln == 0) {
#pragma unroll
      for (int j = 0; j < NS; ++j) k.red[64 + j] = x[j];
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (ln < 2 * NS) {
      const double v = k.red[64 + (ln >> 1)];
      const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
      const unsigned long long half = (ln & 1) ? (bits >> 32) : (bits & 0xffffffffull);
      __hip_atomic_store(base + 16 * k.part() + ln, tag | half, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
    // collect: lane q < 2 NS P waits for word q % (2 NS) of part q / (2 NS),
    // all of them at once, and leaves it in LDS
    unsigned long long* xw = reinterpret_cast<unsigned long long*>(k.red);
    if (ln < 2 * NS * P) {
      unsigned long long* w = base + 16 * (ln / (2 * NS)) + ln % (2 * NS);
      unsigned long long h = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int spins = 0;
      while ((h & ~0xffffffffull) != want) {
        if (++spins > a->spin_limit) {
          // LDS flag for this part; the chain's global word tells the
          // other parts (they check it at every launch start and stop)
          s->xfail = 1;
          __hip_atomic_fetch_or(a->xfail_g + k.c, 1u, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        h = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      xw[ln] = h;
    }
    // the wave's LDS writes land before its reads of them
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    if (ln < NS) {
      double t = 0.0;
      for (int p = 0; p < P; ++p) {
        const unsigned long long lo = xw[p * 2 * NS + 2 * ln], hi = xw[p * 2 * NS + 2 * ln + 1];
        t += __longlong_as_double(static_cast<long long>((lo & 0xffffffffull) | (hi << 32)));
      }
      k.red[64 + ln] = t;
    }
    if (ln == 0) s->xepoch = ep + 1;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NS; ++j) x[j] = k.red[64 + j];
  __syncthreads();
}

// Sum over the chain (fixed order: waves, then wave 0..nw-1, then parts in
// order), the result in every thread.  Called in block-uniform control flow.
__device__ __forceinline__ double block_sum(const ctx& k, double x) {
  x = wave_sum(x);
  if (k.nt > 64) {
    const int w = k.lane >> 6;
    __syncthreads();
    if ((k.lane & 63) == 0) k.red[w] = x;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < (k.nt >> 6); ++i) t += k.red[i];
    __syncthreads();
    x = t;
  }
  if (k.a->parts > 1) {
    double v[1] = {x};
    part_sum<1>(k, v);
    x = v[0];
  }
  return x;
}

// ---- vector helpers (threads stride over D) ----
// Elementwise pass over [0, D): each thread takes its indices kU at a time,
// gathers the inputs of all kU (ld) before it writes any result (st), so
// kU loads are in flight together instead of one load per round trip.
constexpr int kU = 4;
template <class T, class Ld, class St>
__device__ __forceinline__ void vpass(const ctx& k, Ld ld, St st) {
  if (k.s->xfail) return;  // a failed exchange: leave the chain's vectors alone
  const int D = k.d1(), nt = k.nt;
  for (int i0 = k.d0() + k.lane; i0 < D; i0 += kU * nt) {
    T v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * nt;
      if (i < D) v[u] = ld(i);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * nt;
      if (i < D) st(i, v[u]);
    }
  }
}
// Thread-local sum of term(i) over the thread's indices, in index order
// (the loads of kU indices issued together).
template <class Ld>
__device__ __forceinline__ double vsum(const ctx& k, Ld term) {
  const int D = k.d1(), nt = k.nt;
  double acc = 0.0;
  for (int i0 = k.d0() + k.lane; i0 < D; i0 += kU * nt) {
    double t[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * nt;
      t[u] = i < D ? term(i) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
      if (i0 + u * nt < D) acc += t[u];
  }
  return acc;
}

// vpass whose write step also adds to NS thread-local sums, in the index
// order vsum uses (so a fused pass gives the sums of separate passes
// bitwise)
template <int NS, class T, class Ld, class St>
__device__ __forceinline__ void vpass_acc(const ctx& k, double (&acc)[NS], Ld ld, St st) {
  if (k.s->xfail) return;  // a failed exchange: leave the chain's vectors alone
  const int D = k.d1(), nt = k.nt;
  for (int i0 = k.d0() + k.lane; i0 < D; i0 += kU * nt) {
    T v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * nt;
      if (i < D) v[u] = ld(i);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int i = i0 + u * nt;
      if (i < D) st(i, v[u], acc);
    }
  }
}
// block_sum of NS values with one exchange (per value: the same order)
template <int NS>
__device__ __forceinline__ void block_sum_n(const ctx& k, double (&x)[NS]) {
#pragma unroll
  for (int j = 0; j < NS; ++j) x[j] = wave_sum(x[j]);
  if (k.nt > 64) {
    const int w = k.lane >> 6;
    __syncthreads();
    if ((k.lane & 63) == 0)
#pragma unroll
      for (int j = 0; j < NS; ++j) k.red[8 * j + w] = x[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      double t = 0.0;
      for (int i = 0; i < (k.nt >> 6); ++i) t += k.red[8 * j + i];
      x[j] = t;
    }
    __syncthreads();
  }
  if (k.a->parts > 1) part_sum<NS>(k, x);
}

__device__ __forceinline__ void vcopy(const ctx& k, int dst, int src) {
  double* d = k.slot(dst);
  const double* s = k.slot(src);
  vpass<double>(k, [&](int i) { return s[i]; }, [&](int i, double v) { d[i] = v; });
}
__device__ __forceinline__ void vzero(const ctx& k, int dst) {
  double* d = k.slot(dst);
  for (int i = k.d0() + k.lane; i < k.d1(); i += k.nt) d[i] = 0.0;
}
// sum_i p_sharp_i * rho_i with p_sharp = minv * p
__device__ __forceinline__ double dot_sharp(const ctx& k, int p, int rho) {
  const double* m = k.slot(MINV);
  const double* pp = k.slot(p);
  const double* r = k.slot(rho);
  const double acc = vsum(k, [&](int i) { return (m[i] * pp[i]) * r[i]; });
  return block_sum(k, acc);
}
// sum_i p_sharp_i * (r1_i + r2_i)
__device__ __forceinline__ double dot_sharp2(const ctx& k, int p, int r1, int r2) {
  const double* m = k.slot(MINV);
  const double* pp = k.slot(p);
  const double* a = k.slot(r1);
  const double* b = k.slot(r2);
  const double acc = vsum(k, [&](int i) { return (m[i] * pp[i]) * (a[i] + b[i]); });
  return block_sum(k, acc);
}
__device__ __forceinline__ double kinetic(const ctx& k, int p) {
  const double* m = k.slot(MINV);
  const double* pp = k.slot(p);
  const double acc = vsum(k, [&](int i) { return pp[i] * (m[i] * pp[i]); });
  return 0.5 * block_sum(k, acc);
}

// ---- RNG: the chain's Philox stream, one variate per counter value ----
// Called in block-uniform control flow: every thread draws the same value,
// then thread 0 advances the counter.
__device__ double uniform(const ctx& k) {
  const uint64_t ctr = k.s->rng;
  const double u = stanx::philox_uniform(k.a->seed, k.chain_id, ctr);
  __syncthreads();
  if (k.lane == 0) k.s->rng = ctr + 1;
  __syncthreads();
  return u;
}
// diag_e_metric::sample_p: p_i = z_i / sqrt(minv_i), variate i of the call
// from counter base + i
__device__ void sample_p(const ctx& k) {
  double* p = k.slot(ZP);
  const double* m = k.slot(MINV);
  const uint64_t base = k.s->rng;
  for (int i = k.d0() + k.lane; i < k.d1(); i += k.nt)
    p[i] = stanx::philox_normal(k.a->seed, k.chain_id, base + i) / sqrt(m[i]);
  __syncthreads();
  if (k.lane == 0) k.s->rng = base + k.a->D;
  __syncthreads();
}

__device__ double log1p_exp(double a) { return a > 0.0 ? a + log1p(exp(-a)) : log1p(exp(a)); }
__device__ double log_sum_exp(double a, double b) {
  if (a == -kInf) return b;
  if (a == kInf && b == kInf) return kInf;
  if (a > b) return a + log1p_exp(b - a);
  return b + log1p_exp(a - b);
}

__device__ double& lsw_ref(const ctx& k, int id) {
  if (id < 0) return k.s->lsw_sub;
  return (id & 1) ? k.s->lsw_final[id >> 1] : k.s->lsw_init[id >> 1];
}
__device__ int prop_slot(int id) { return id < 0 ? PQ : (id == 0 ? SQ : ZPF(id)); }
__device__ double& prop_V(const ctx& k, int id) {
  return id < 0 ? k.s->V_prop : (id == 0 ? k.s->V_samp : k.s->V_zpf[id]);
}
__device__ double& prop_T(const ctx& k, int id) {
  return id < 0 ? k.s->T_prop : (id == 0 ? k.s->T_samp : k.s->T_zpf[id]);
}
__device__ void copy_prop(const ctx& k, int dst, int src) {
  vcopy(k, prop_slot(dst), prop_slot(src));
  if (k.lane == 0) {
    prop_V(k, dst) = prop_V(k, src);
    prop_T(k, dst) = prop_T(k, src);
  }
}

// ---- point copies (q, p, g, V) ----
__device__ void copy_point(const ctx& k, int dq, int sq, double& dV, double sV) {
  vcopy(k, dq, sq);
  vcopy(k, dq + 1, sq + 1);
  vcopy(k, dq + 2, sq + 2);
  if (k.lane == 0) dV = sV;
}

// ---- gradient request / consumption ----
__device__ void request_gradient(const ctx& k, int next_phase) {
  double* th = k.a->theta + static_cast<int64_t>(k.c) * k.a->D;
  const double* q = k.slot(ZQ);
  vpass<double>(k, [&](int i) { return q[i]; }, [&](int i, double v) { th[i] = v; });
  if (k.lane == 0) {
    k.s->phase = next_phase;
    k.s->pending = 1;
    k.s->n_grad++;
  }
}
// base_hamiltonian::update_potential_gradient: V = -lp, g = -grad; on a
// domain error V = +inf and g = -g_old (the reference flips the stale g)
// With lf, the second half of the leapfrog (p -= eps/2 g) follows in the
// same pass: the phases that requested this gradient (PH_LEAF, PH_SS_B)
// begin with it.
__device__ void consume_gradient(const ctx& k, bool lf, double lf_eps) {
  const double* o = k.a->out + static_cast<int64_t>(k.c) * k.a->out_stride;
  const bool ok = o[k.a->D + 1] == 0.0 && isfinite(o[0]);
  double* g = k.slot(ZG);
  if (lf) {
    double* p = k.slot(ZP);
    const double he = 0.5 * lf_eps;
    struct gp { double g, p; };
    vpass<gp>(k, [&](int i) { return gp{ok ? o[1 + i] : g[i], p[i]}; },
              [&](int i, const gp& v) {
                const double gn = -v.g;
                g[i] = gn;
                p[i] = v.p - he * gn;
              });
  } else {
    vpass<double>(k, [&](int i) { return ok ? o[1 + i] : g[i]; }, [&](int i, double v) { g[i] = -v; });
  }
  if (k.lane == 0) {
    k.s->V_z = ok ? -o[0] : kInf;
    k.s->last_lp = o[0];
    k.s->last_ok = ok ? 1 : 0;
  }
}

// the leapfrog's begin half (expl_leapfrog: p -= e/2 g; q += e minv p)
// and the gradient request at the new q, in one pass
__device__ void leapfrog_request(const ctx& k, double eps, int next_phase) {
  double* p = k.slot(ZP);
  double* q = k.slot(ZQ);
  const double* g = k.slot(ZG);
  const double* m = k.slot(MINV);
  double* th = k.a->theta + static_cast<int64_t>(k.c) * k.a->D;
  const double he = 0.5 * eps;
  struct pgqm { double p, g, q, m; };
  vpass<pgqm>(k, [&](int i) { return pgqm{p[i], g[i], q[i], m[i]}; },
              [&](int i, const pgqm& v) {
                const double pn = v.p - he * v.g;
                const double qn = v.q + eps * (v.m * pn);
                p[i] = pn;
                q[i] = qn;
                th[i] = qn;
              });
  if (k.lane == 0) {
    k.s->phase = next_phase;
    k.s->pending = 1;
    k.s->n_grad++;
  }
}
__device__ void leapfrog_end(const ctx& k, double eps) {
  double* p = k.slot(ZP);
  const double* g = k.slot(ZG);
  const double he = 0.5 * eps;
  struct pg { double p, g; };
  vpass<pg>(k, [&](int i) { return pg{p[i], g[i]}; },
            [&](int i, const pg& v) { p[i] = v.p - he * v.g; });
}

__device__ dest init_dest(const dest& d, int l) {
  return dest{d.zprop, d.pbeg, PIE(l), RI(l), 2 * l};
}
__device__ dest final_dest(const dest& d, int l) {
  return dest{l, PFBG(l), d.pend, RF(l), 2 * l + 1};
}
__device__ dest top_dest(const ctx& k) {
  return k.s->fwd ? dest{-1, PFB, FP, RHOF, -1} : dest{-1, PBF, BP, RHOB, -1};
}
__device__ dest dest_of(const ctx& k, int level) {
  dest d = top_dest(k);
  for (int l = k.s->tree_depth; l > level; --l)
    d = k.s->stage[l] == 0 ? init_dest(d, l) : final_dest(d, l);
  return d;
}

// base case of build_tree, fused with the leapfrog's second half.  One
// pass: g = -grad (or the flipped stale g after a domain error), p -= e/2 g,
// kinetic energy, proposal position, rho, subtree ends.  The kinetic energy
// is returned in T.
//
// Every rho of the tree (rho_init, rho_final of each level, rho_fwd /
// rho_bck of the top-level subtree) starts as a zero vector and receives
// exactly one "rho += x" (base_nuts.hpp: the base case adds p, a merged
// subtree adds rho_init + rho_final, each into its caller's rho), so the
// slots are never read before that write, which stores 0 + x (bitwise the
// sum into a cleared slot) and need no clearing.  Likewise the final
// half's proposal, z_propose_final = z at its start, is overwritten by its
// first leaf before any merge reads it, so its start is not stored.
//
// With ls > 0 (the first level whose initial half this leaf completes if
// the merges below it hold; spec) the tree, if it goes on, goes on from
// this point in the same direction (the final half of level ls starts
// here), so the same pass also takes the next leapfrog's begin half:
// p' = p - e/2 g into NP and q' = q + e minv p' into theta, the gradient
// request position; the final half's start then requests that gradient
// without another pass, and the next leaf reads its point as q = theta,
// p = NP (znext).  The arithmetic is that of the separate pass.  A spec
// leaf's own point is not stored at ZQ, ZP, ZG (its p and q are kept in
// the subtree-end and proposal slots): the point is the current one again
// only if the tree stops, and then its trajectory ends are not read again
// (the transition ends).  Its g is read next only by the next leaf after a
// domain error (the flipped stale g), and that leaf is divergent (V = inf):
// the tree stops there and none of its values is read again, so a stale g
// from an earlier leaf serves.  A leaf without spec stores the whole point
// (q, p, g) at ZQ, ZP, ZG.
__device__ bool leaf(const ctx& k, const dest& d, double lf_eps, bool ok, int ls, double& T) {
  chain_state* s = k.s;
  double acc[1] = {0.0};
  {
    const double* o = k.a->out + static_cast<int64_t>(k.c) * k.a->out_stride;
    const bool zn = s->znext != 0;
    const bool spec = ls > 0;
    double* th = k.a->theta + static_cast<int64_t>(k.c) * k.a->D;
    double* g = k.slot(ZG);
    double* p = k.slot(ZP);
    double* zq = k.slot(ZQ);
    double* np = k.slot(NP);
    const double* pin = zn ? np : p;
    const double* qin = zn ? th : zq;
    const double* m = k.slot(MINV);
    double* prop = k.slot(prop_slot(d.zprop));
    double* r = k.slot(d.rho);
    double* pb = k.slot(d.pbeg);
    double* pe = k.slot(d.pend);
    // below the top (tree_depth >= 1) the leaf is the initial or the final
    // half of a level-1 subtree, whose merge reads its p from the subtree's
    // begin or end slot and forms its rho as 0 + p (see merge): the leaf
    // stores p once and no rho
    const bool top = s->tree_depth == 0;
    const bool wb = top || s->stage[1] == 0;
    const bool we = top || s->stage[1] != 0;
    const double he = 0.5 * lf_eps;
    struct lf { double g, p, m, q; };
    vpass_acc<1, lf>(k, acc, [&](int i) { return lf{ok ? o[1 + i] : g[i], pin[i], m[i], qin[i]}; },
                     [&](int i, const lf& v, double (&ac)[1]) {
                       const double gn = -v.g;
                       const double pn = v.p - he * gn;
                       ac[0] += pn * (v.m * pn);
                       prop[i] = v.q;
                       if (top) r[i] = 0.0 + pn;
                       if (wb) pb[i] = pn;
                       if (we) pe[i] = pn;
                       if (spec) {
                         const double p2 = pn - he * gn;
                         np[i] = p2;
                         th[i] = v.q + lf_eps * (v.m * p2);
                       } else {
                         g[i] = gn;
                         p[i] = pn;
                         if (zn) zq[i] = v.q;
                       }
                     });
  }
  block_sum_n<1>(k, acc);
  T = 0.5 * acc[0];
  double h = T + s->V_z;
  if (isnan(h)) h = kInf;
  const double H0 = s->H0;
  // every wave has read znext (block_sum_n's barriers, or one wave)
  if (k.lane == 0) {
    s->znext = 0;
    s->spec = ls > 0 ? 1 : 0;
    s->n_leapfrog++;
    if ((h - H0) > k.a->max_deltaH) s->divergent = 1;
    double& lw = lsw_ref(k, d.lsw);
    lw = log_sum_exp(lw, H0 - h);
    if (H0 - h > 0)
      s->smp += 1;
    else
      s->smp += exp(H0 - h);
    prop_V(k, d.zprop) = s->V_z;
    prop_T(k, d.zprop) = T;
  }
  __syncthreads();
  return !s->divergent;
}

// The vector pass of merge().  At level 1 both halves are single leaves:
// the initial half's p (its begin and end) is at the merged subtree's begin
// slot, the final half's p at its end slot, and their rho_init and
// rho_final, 0 + p (the leaf's one contribution), are formed here (L1).
template <bool L1>
__device__ void merge_pass(const ctx& k, int l, const dest& d, bool take, double (&acc)[6]) {
  const double* m = k.slot(MINV);
  const double* pend = k.slot(d.pend);
  const double* pbeg = k.slot(d.pbeg);
  const double* psrc = k.slot(prop_slot(l));
  double* pdst = k.slot(prop_slot(d.zprop));
  double* r = k.slot(d.rho);
  auto body = [&](int i, double ri, double rf, double mv, double pe, double pb, double pfbg,
                  double pie, double prop, double (&ac)[6]) {
    if (take) pdst[i] = prop;
    const double tv = ri + rf;
    r[i] = 0.0 + tv;
    ac[0] += (mv * pe) * tv;
    ac[1] += (mv * pb) * tv;
    ac[2] += (mv * pfbg) * (ri + pfbg);
    ac[3] += (mv * pb) * (ri + pfbg);
    ac[4] += (mv * pe) * (rf + pie);
    ac[5] += (mv * pie) * (rf + pie);
  };
  if constexpr (L1) {
    struct mg { double m, pend, pbeg, prop; };
    vpass_acc<6, mg>(
        k, acc, [&](int i) { return mg{m[i], pend[i], pbeg[i], take ? psrc[i] : 0.0}; },
        [&](int i, const mg& v, double (&ac)[6]) {
          body(i, 0.0 + v.pbeg, 0.0 + v.pend, v.m, v.pend, v.pbeg, v.pend, v.pbeg, v.prop, ac);
        });
  } else {
    const double* ri = k.slot(RI(l));
    const double* rf = k.slot(RF(l));
    const double* pfbg = k.slot(PFBG(l));
    const double* pie = k.slot(PIE(l));
    struct mg { double ri, rf, m, pend, pbeg, pfbg, pie, prop; };
    vpass_acc<6, mg>(
        k, acc,
        [&](int i) {
          return mg{ri[i], rf[i], m[i], pend[i], pbeg[i], pfbg[i], pie[i], take ? psrc[i] : 0.0};
        },
        [&](int i, const mg& v, double (&ac)[6]) {
          body(i, v.ri, v.rf, v.m, v.pend, v.pbeg, v.pfbg, v.pie, v.prop, ac);
        });
  }
}

// merge of the two halves of level l into d (base_nuts.hpp:321-351)
__device__ bool merge(const ctx& k, int l, const dest& d) {
  chain_state* s = k.s;
  const double lsw_sub = log_sum_exp(s->lsw_init[l], s->lsw_final[l]);
  double& lw = lsw_ref(k, d.lsw);
  const double new_lw = log_sum_exp(lw, lsw_sub);
  bool take;
  if (s->lsw_final[l] > lsw_sub) {
    take = true;
  } else {
    const double accept_prob = exp(s->lsw_final[l] - lsw_sub);
    take = uniform(k) < accept_prob;
  }
  __syncthreads();
  if (k.lane == 0) {
    lw = new_lw;
    if (take) {
      prop_V(k, d.zprop) = prop_V(k, l);
      prop_T(k, d.zprop) = prop_T(k, l);
    }
  }
  // one pass: the proposal copy, rho_subtree = rho_init + rho_final,
  // rho(d) = 0 + rho_subtree (its one contribution), and the six sums of
  // the three U-turn checks (p_sharp = minv p): merged (pend, pbeg against
  // rho_subtree), left join (pfbg, pbeg against rho_init + pfbg), right
  // join (pend, pie against rho_final + pie)
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (l == 1)
    merge_pass<true>(k, l, d, take, acc);
  else
    merge_pass<false>(k, l, d, take, acc);
  block_sum_n<6>(k, acc);
  bool persist = acc[0] > 0 && acc[1] > 0;
  persist &= acc[2] > 0 && acc[3] > 0;
  persist &= acc[4] > 0 && acc[5] > 0;
  __syncthreads();
  return persist;
}

// dual averaging (stepsize_adaptation::learn_stepsize)
__device__ void learn_stepsize(const ctx& k, double stat) {
  chain_state* s = k.s;
  const batch_args* a = k.a;
  s->sa_counter += 1;
  stat = stat > 1 ? 1 : stat;
  const double eta = 1.0 / (s->sa_counter + a->t0);
  s->s_bar = (1.0 - eta) * s->s_bar + eta * (a->delta - stat);
  const double x = s->mu - s->s_bar * sqrt(s->sa_counter) / a->gamma;
  const double x_eta = pow(s->sa_counter, -a->kappa);
  s->x_bar = (1.0 - x_eta) * s->x_bar + x_eta * x;
  s->nom_eps = exp(x);
}

// windowed variance adaptation over the run's window plan
// (stanx/mcmc/warmup_plan.hpp): does this transition's draw feed the
// running variance?
__device__ bool in_adapt_window(const ctx& k) {
  return stanx::mcmc::plan_feeds(k.a->wplan, k.s->w_counter);
}
// Welford update of the variance estimate with the draw (at ZQ); the
// end-of-transition pass may already have applied it (welford_done)
__device__ int learn_variance(const ctx& k, bool welford_done) {
  chain_state* s = k.s;
  if (in_adapt_window(k)) {
    const double n = s->w_n + 1.0;
    if (!welford_done) {
      double* m = k.slot(WM);
      double* m2 = k.slot(WM2);
      const double* q = k.slot(ZQ);
      struct qmm { double q, m, m2; };
      vpass<qmm>(k, [&](int i) { return qmm{q[i], m[i], m2[i]}; },
                 [&](int i, const qmm& v) {
                   const double delta = v.q - v.m;
                   const double mn = v.m + delta / n;
                   m[i] = mn;
                   m2[i] = v.m2 + (v.q - mn) * delta;
                 });
    }
    __syncthreads();
    if (k.lane == 0) s->w_n = n;
  }
  __syncthreads();
  const bool end = stanx::mcmc::plan_closes(k.a->wplan, s->w_cursor, s->w_counter);
  if (!end) {
    __syncthreads();
    if (k.lane == 0) s->w_counter++;
    return 0;
  }
  const double n = s->w_n;
  double* minv = k.slot(MINV);
  double* m = k.slot(WM);
  double* m2 = k.slot(WM2);
  int finite = 1;
  struct vm { double minv, m2; };
  vpass<vm>(k, [&](int i) { return vm{minv[i], m2[i]}; },
            [&](int i, const vm& v) {
              double var = v.minv;
              if (n > 1) var = v.m2 / (n - 1.0);
              var = (n / (n + 5.0)) * var + 1e-3 * (5.0 / (n + 5.0)) * 1.0;
              if (!isfinite(var)) finite = 0;
              minv[i] = var;
              m[i] = 0.0;
              m2[i] = 0.0;
            });
  int all_finite = __syncthreads_and(finite);
  __syncthreads();
  if (k.a->parts > 1) {
    double nbad[1] = {all_finite ? 0.0 : 1.0};
    part_sum<1>(k, nbad);
    all_finite = nbad[0] == 0.0;
  }
  if (k.lane == 0) {
    s->w_cursor++;
    s->w_n = 0;
    s->w_counter++;
  }
  return all_finite ? 1 : -1;
}

__device__ void finish(const ctx& k, int status) {
  if (k.lane == 0) {
    k.s->status = status;
    k.s->phase = PH_DONE;
    k.s->pending = 0;
    if (k.part() == 0) atomicAdd(k.a->done_count, 1u);
  }
}

__device__ void begin_stepsize_heuristic(const ctx& k, int ret_phase) {
  chain_state* s = k.s;
  if (s->nom_eps == 0 || s->nom_eps > 1e7 || isnan(s->nom_eps)) {
    if (k.lane == 0) s->phase = ret_phase;
    return;
  }
  copy_point(k, IQ, ZQ, s->V_init, s->V_z);
  __syncthreads();
  if (k.lane == 0) {
    s->ss_ret = ret_phase;
    s->ss_iter = 0;
  }
  sample_p(k);
  __syncthreads();
  request_gradient(k, PH_SS_A);
}

}  // namespace

__global__ __launch_bounds__(512) void nuts_step(batch_args a) {
  __shared__ chain_state s;
  __shared__ batch_args sa;  // helpers read the arguments through ctx::a
  __shared__ double red[72];  // block_sum_n: 8 waves x up to 6 values; part_sum: 64 words + 8 sums
  __shared__ int slice[3];    // d0, d1, part
  // chain and part of this workgroup; with C a multiple of 8 the parts of
  // a chain are 8 block ids apart, i.e. on one XCD (shared L2)
  const int P = a.parts;
  const int bid = blockIdx.x;
  int c, part;
  if (P > 1 && a.C % 8 == 0) {
    part = (bid >> 3) % P;
    c = (bid & 7) + 8 * ((bid >> 3) / P);
  } else {
    part = bid % P;
    c = bid / P;
  }
  // block-uniform: keep them (and every address derived from them) scalar
  c = __builtin_amdgcn_readfirstlane(c);
  part = __builtin_amdgcn_readfirstlane(part);
  const int lane = threadIdx.x;
  const int nt = blockDim.x;
  {
    const int* src = reinterpret_cast<const int*>(a.states + c);
    int* dst = reinterpret_cast<int*>(&s);
    for (int i = lane; i < static_cast<int>(sizeof(chain_state) / 4); i += nt) dst[i] = src[i];
  }
  if (lane == 0) {
    sa = a;
    // slice of the part: whole multiples of 64 elements, the last ends at D
    const int span = ((a.D + P - 1) / P + 63) & ~63;
    const int d0 = part * span < a.D ? part * span : a.D;
    slice[0] = d0;
    slice[1] = d0 + span < a.D ? d0 + span : a.D;
    slice[2] = part;
  }
  __syncthreads();
  ctx k{&sa, &s, a.vecs + static_cast<int64_t>(c) * a.nslot * a.Dpad, c, lane, nt, red,
        a.chain0 + static_cast<uint32_t>(c), slice};
  if (s.phase == PH_DONE) return;
  // another part of this chain timed out in an exchange: stop here, before
  // any vector is touched
  if (P > 1) {
    __shared__ int stop;
    if (lane == 0)
      stop = __hip_atomic_load(a.xfail_g + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    __syncthreads();
    if (stop) {
      finish(k, 70);
      __syncthreads();
      if (part == 0) {
        const int* src = reinterpret_cast<const int*>(&s);
        int* dst = reinterpret_cast<int*>(a.states + c);
        for (int i = lane; i < static_cast<int>(sizeof(chain_state) / 4); i += nt) dst[i] = src[i];
      }
      return;
    }
  }
  // a PH_LEAF gradient is consumed inside leaf(), fused with its pass
  bool leaf_ok = false;
  if (s.pending) {
    if (s.phase == PH_LEAF) {
      const double* o = a.out + static_cast<int64_t>(c) * a.out_stride;
      leaf_ok = o[a.D + 1] == 0.0 && isfinite(o[0]);
      __syncthreads();
      if (lane == 0) {
        s.V_z = leaf_ok ? -o[0] : kInf;
        s.last_lp = o[0];
        s.last_ok = leaf_ok ? 1 : 0;
      }
    } else {
      // PH_SS_B begins with the leapfrog's second half
      const bool lf = s.phase == PH_SS_B;
      consume_gradient(k, lf, s.nom_eps);
    }
    __syncthreads();
    if (lane == 0) s.pending = 0;
  }
  __syncthreads();
  // advance until a gradient is requested, the chain finishes, or it has
  // taken phase_budget phase turns in this launch: it then continues where
  // it stopped in the next launch (the step between evaluates the same
  // point again, unread).  The bound keeps the launch from waiting on the
  // one chain whose leaf closes a deep subtree or a transition while the
  // others take one leapfrog each; the arithmetic and the random stream
  // are those of an unbounded run.  A final half's start whose leapfrog
  // the leaf's pass already took (no vector pass; it ends the launch with
  // its gradient request) is taken whatever the count.
  for (int turn = 0; !s.pending && s.phase != PH_DONE && !s.xfail; ++turn) {
    const int ph = s.phase;
    if (turn >= a.phase_budget &&
        !(ph == PH_MERGE && s.mvalid && s.mlevel <= s.tree_depth && s.stage[s.mlevel] == 0 &&
          s.spec))
      break;
    __syncthreads();
    switch (ph) {
      case PH_INIT_TRY: {
        // uniform(-R, R) inits (random_var_context), or the given init
        double* q = k.slot(ZQ);
        if (a.inits) {
          const double* src = a.inits + static_cast<int64_t>(c) * a.D;
          for (int i = k.d0() + lane; i < k.d1(); i += k.nt) q[i] = src[i];
        } else {
          const uint64_t base = s.rng;
          for (int i = k.d0() + lane; i < k.d1(); i += k.nt)
            q[i] = a.init_radius == 0.0
                       ? 0.0
                       : -a.init_radius + stanx::philox_uniform(a.seed, k.chain_id, base + i) *
                                              2.0 * a.init_radius;
          __syncthreads();
          if (lane == 0 && a.init_radius != 0.0) s.rng = base + a.D;
        }
        __syncthreads();
        request_gradient(k, PH_INIT_CHECK);
        break;
      }
      case PH_INIT_CHECK: {
        const double* g = k.slot(ZG);
        double gs = vsum(k, [&](int i) { return g[i]; });
        gs = block_sum(k, gs);
        const bool ok = s.last_ok && isfinite(s.last_lp) && isfinite(gs);
        __syncthreads();
        if (ok) {
          // run_adaptive_sampler: engage adaptation, step-size heuristic
          if (lane == 0) s.adapt_flag = 1;
          begin_stepsize_heuristic(k, PH_TRANS_BEGIN);
        } else {
          if (lane == 0) s.init_tries++;
          __syncthreads();
          if (s.init_tries >= a.max_init_tries) {
            finish(k, 78);
          } else if (lane == 0) {
            s.phase = PH_INIT_TRY;
          }
        }
        break;
      }
      case PH_SS_A: {  // heuristic: gradient at z with fresh p
        const double T = kinetic(k, ZP);
        if (lane == 0) s.ss_H0 = T + s.V_z;
        leapfrog_request(k, s.nom_eps, PH_SS_B);
        break;
      }
      case PH_SS_B: {  // the leapfrog's second half ran with the gradient
        const double T = kinetic(k, ZP);
        double h = T + s.V_z;
        if (isnan(h)) h = kInf;
        const double dH = s.ss_H0 - h;
        bool done = false, fail = false;
        double eps = s.nom_eps;
        int dir = s.ss_dir;
        if (s.ss_iter == 0) {
          dir = dH > kLog08 ? 1 : -1;
        } else if ((dir == 1) && !(dH > kLog08)) {
          done = true;
        } else if ((dir == -1) && !(dH < kLog08)) {
          done = true;
        } else {
          eps = dir == 1 ? 2.0 * eps : 0.5 * eps;
          if (eps > 1e7 || eps == 0) fail = true;
        }
        __syncthreads();
        if (lane == 0) {
          s.ss_dir = dir;
          s.ss_iter = 1;
          s.nom_eps = eps;
        }
        // restore z_init
        copy_point(k, ZQ, IQ, s.V_z, s.V_init);
        __syncthreads();
        if (fail) {
          // the reference's initial heuristic failure is logged and the
          // service returns OK without draws; later failures are SOFTWARE
          finish(k, s.ss_ret == PH_TRANS_BEGIN ? 0 : 70);
        } else if (done) {
          if (lane == 0) s.phase = s.ss_ret;
        } else {
          sample_p(k);
          __syncthreads();
          request_gradient(k, PH_SS_A);
        }
        break;
      }
      case PH_POST_ADAPT: {
        if (lane == 0) {
          s.mu = log(10 * s.nom_eps);
          s.sa_counter = 0;
          s.s_bar = 0;
          s.x_bar = 0;
          s.iter++;
          s.phase = PH_TRANS_BEGIN;
        }
        break;
      }
      case PH_TRANS_BEGIN: {
        if (s.iter >= s.iter_end) {
          finish(k, 0);
          break;
        }
        // every wave reads the condition before lane 0 changes adapt_flag
        const bool end_warmup = s.iter == a.num_warmup && s.adapt_flag;
        __syncthreads();
        if (end_warmup) {
          // disengage_adaptation: complete_adaptation sets eps = exp(x_bar)
          if (lane == 0) {
            s.adapt_flag = 0;
            s.nom_eps = exp(s.x_bar);
          }
          __syncthreads();
        }
        // sample_stepsize, seed(q) (z.q already holds the last sample), sample_p
        double eps = s.nom_eps;
        if (a.jitter) eps *= 1.0 + a.jitter * (2.0 * uniform(k) - 1.0);
        __syncthreads();
        if (lane == 0) s.eps = eps;
        // one pass: sample_p and the gradient request at z
        {
          double* p = k.slot(ZP);
          const double* m = k.slot(MINV);
          const double* q = k.slot(ZQ);
          double* th = a.theta + static_cast<int64_t>(c) * a.D;
          const uint64_t base = s.rng;
          for (int i = k.d0() + lane; i < k.d1(); i += nt) {
            p[i] = stanx::philox_normal(a.seed, k.chain_id, base + i) / sqrt(m[i]);
            th[i] = q[i];
          }
        }
        __syncthreads();
        if (lane == 0) {
          s.rng += a.D;
          s.phase = PH_TRANS_INIT;
          s.pending = 1;
          s.n_grad++;
        }
        break;
      }
      case PH_TRANS_INIT: {
        // one pass: kinetic energy at z; both trajectory ends, the sample
        // and proposal positions, the two subtree-end momenta and rho from z
        double kacc[1] = {0.0};
        {
          const double* q = k.slot(ZQ);
          const double* p = k.slot(ZP);
          const double* g = k.slot(ZG);
          const double* m = k.slot(MINV);
          double* fq = k.slot(FQ);
          double* fp = k.slot(FP);
          double* fg = k.slot(FG);
          double* bq = k.slot(BQ);
          double* bp = k.slot(BP);
          double* bg = k.slot(BG);
          double* sq = k.slot(SQ);
          double* pq = k.slot(PQ);
          double* pfb = k.slot(PFB);
          double* pbf = k.slot(PBF);
          double* rho = k.slot(RHO);
          struct zm { double q, p, g, m; };
          vpass_acc<1, zm>(k, kacc, [&](int i) { return zm{q[i], p[i], g[i], m[i]}; },
                           [&](int i, const zm& v, double (&ac)[1]) {
                             ac[0] += v.p * (v.m * v.p);
                             fq[i] = v.q;
                             fp[i] = v.p;
                             fg[i] = v.g;
                             bq[i] = v.q;
                             bp[i] = v.p;
                             bg[i] = v.g;
                             sq[i] = v.q;
                             pq[i] = v.q;
                             pfb[i] = v.p;
                             pbf[i] = v.p;
                             rho[i] = v.p;
                           });
        }
        block_sum_n<1>(k, kacc);
        const double T = 0.5 * kacc[0];
        __syncthreads();
        if (lane == 0) {
          s.V_fwd = s.V_z;
          s.V_bck = s.V_z;
          s.V_samp = s.V_z;
          s.T_samp = T;
          s.V_prop = s.V_z;
          s.T_prop = T;
          s.lsw = 0;
          s.H0 = T + s.V_z;
          s.n_leapfrog = 0;
          s.smp = 0;
          s.depth = 0;
          s.divergent = 0;
          s.phase = PH_DOUBLING;
        }
        break;
      }
      case PH_DOUBLING: {
        if (s.depth >= a.max_depth) {
          if (lane == 0) s.phase = PH_TRANS_END;
          break;
        }
        const bool fwd = uniform(k) > 0.5;
        __syncthreads();
        // one pass: z = the chosen end; rho of the other side = rho; the
        // other side's end momentum; then the first leapfrog's begin half
        // and its gradient request (leapfrog_request).  The new subtree's
        // rho and every level's rho_init are written by their one
        // contribution (see leaf), not cleared here.
        const int td = s.depth;
        const double eps = (fwd ? 1.0 : -1.0) * s.eps;
        {
          const double* eq = k.slot(fwd ? FQ : BQ);
          const double* ep = k.slot(fwd ? FP : BP);
          const double* eg = k.slot(fwd ? FG : BG);
          const double* m = k.slot(MINV);
          const double* rho = k.slot(RHO);
          double* zq = k.slot(ZQ);
          double* zp = k.slot(ZP);
          double* zg = k.slot(ZG);
          double* rother = k.slot(fwd ? RHOB : RHOF);
          double* pother = k.slot(fwd ? PBF : PFB);
          double* th = a.theta + static_cast<int64_t>(c) * a.D;
          const double he = 0.5 * eps;
          struct ez { double q, p, g, m, r; };
          vpass<ez>(k, [&](int i) { return ez{eq[i], ep[i], eg[i], m[i], rho[i]}; },
                    [&](int i, const ez& v) {
                      const double pn = v.p - he * v.g;
                      const double qn = v.q + eps * (v.m * pn);
                      zq[i] = qn;
                      zp[i] = pn;
                      zg[i] = v.g;
                      th[i] = qn;
                      rother[i] = v.r;
                      pother[i] = v.p;
                    });
        }
        __syncthreads();
        if (lane == 0) {
          s.V_z = fwd ? s.V_fwd : s.V_bck;
          s.fwd = fwd ? 1 : 0;
          s.lsw_sub = -kInf;
          s.tree_depth = td;
          for (int l = 1; l <= td; ++l) {
            s.stage[l] = 0;
            s.lsw_init[l] = -kInf;
          }
          s.phase = PH_LEAF;
          s.pending = 1;
          s.n_grad++;
        }
        break;
      }
      case PH_LEAF: {  // the leapfrog's second half ran with the gradient
        int ls = 0;  // first level whose initial half is still open
        for (int l = 1; l <= s.tree_depth && ls == 0; ++l)
          if (s.stage[l] == 0) ls = l;
        double T_leaf = 0.0;
        const bool valid = leaf(k, dest_of(k, 0), (s.fwd ? 1.0 : -1.0) * s.eps, leaf_ok, ls, T_leaf);
        if (lane == 0) {
          s.T_leaf = T_leaf;
          s.mlevel = 1;
          s.mvalid = valid ? 1 : 0;
          s.phase = PH_MERGE;
        }
        break;
      }
      case PH_MERGE: {  // the ladder above a leaf, one level per loop turn
        const int l = s.mlevel;
        const bool valid = s.mvalid != 0;
        if (valid && l <= s.tree_depth) {
          if (s.stage[l] == 0) {
            // initial half done (l is the leaf's ls): the final half starts
            // from the current point (its proposal energies are set here as
            // the reference's z_propose_final = z; the first leaf of the
            // half overwrites them and the position, which is not stored)
            __syncthreads();
            if (lane == 0) {
              s.stage[l] = 1;
              s.V_zpf[l] = s.V_z;
              s.T_zpf[l] = s.T_leaf;
              s.lsw_final[l] = -kInf;
            }
            __syncthreads();
            // the leaf's pass took the begin half of the next leapfrog (NP,
            // theta): request the gradient there.  Otherwise one pass: the
            // leapfrog's begin half with its gradient request
            // (leapfrog_request)
            const bool spec = s.spec != 0;
            if (!spec) {
              const double eps = (s.fwd ? 1.0 : -1.0) * s.eps;
              double* p = k.slot(ZP);
              double* q = k.slot(ZQ);
              const double* g = k.slot(ZG);
              const double* m = k.slot(MINV);
              double* th = a.theta + static_cast<int64_t>(c) * a.D;
              const double he = 0.5 * eps;
              struct pgqm { double p, g, q, m; };
              vpass<pgqm>(k, [&](int i) { return pgqm{p[i], g[i], q[i], m[i]}; },
                          [&](int i, const pgqm& v) {
                            const double pn = v.p - he * v.g;
                            const double qn = v.q + eps * (v.m * pn);
                            p[i] = pn;
                            q[i] = qn;
                            th[i] = qn;
                          });
            }
            __syncthreads();
            if (lane == 0) {
              for (int mm = 1; mm < l; ++mm) {
                s.stage[mm] = 0;
                s.lsw_init[mm] = -kInf;
              }
              s.znext = spec ? 1 : 0;
              s.spec = 0;
              s.phase = PH_LEAF;
              s.pending = 1;
              s.n_grad++;
            }
            break;
          }
          const bool v = merge(k, l, dest_of(k, l));
          __syncthreads();
          if (lane == 0) {
            s.mvalid = v ? 1 : 0;
            s.mlevel = l + 1;
          }
          break;
        }
        // the subtree of depth tree_depth is complete (or invalid)
        if (!valid) {
          // the transition ends here: the reference's last z_fwd / z_bck
          // assignment is not read again, so the trajectory end's vectors
          // are left as they are (the next transition sets both ends)
          __syncthreads();
          if (lane == 0) {
            if (s.fwd)
              s.V_fwd = s.V_z;
            else
              s.V_bck = s.V_z;
          }
          __syncthreads();
          if (lane == 0) s.phase = PH_TRANS_END;
          break;
        }
        bool take;
        if (s.lsw_sub > s.lsw) {
          take = true;
        } else {
          take = uniform(k) < exp(s.lsw_sub - s.lsw);
        }
        __syncthreads();
        const bool fwd = s.fwd != 0;
        if (lane == 0) {
          if (fwd)
            s.V_fwd = s.V_z;
          else
            s.V_bck = s.V_z;
          if (take) {
            s.V_samp = s.V_prop;
            s.T_samp = s.T_prop;
          }
        }
        // one pass: the new trajectory end (z), the sample copy, rho =
        // rho_bck + rho_fwd and the six sums of the top-level U-turn checks
        double acc[6] = {0, 0, 0, 0, 0, 0};
        {
          const double* zq = k.slot(ZQ);
          const double* zp = k.slot(ZP);
          const double* zg = k.slot(ZG);
          const double* pq = k.slot(PQ);
          const double* rb = k.slot(RHOB);
          const double* rf = k.slot(RHOF);
          const double* m = k.slot(MINV);
          const double* oth = k.slot(fwd ? BP : FP);
          const double* pfb = k.slot(PFB);
          const double* pbf = k.slot(PBF);
          double* eq = k.slot(fwd ? FQ : BQ);
          double* sq = k.slot(SQ);
          double* rho = k.slot(RHO);
          struct tl { double zq, zp, zg, pq, rb, rf, m, oth, pfb, pbf; };
          vpass_acc<6, tl>(
              k, acc,
              [&](int i) {
                return tl{zq[i], zp[i], zg[i], take ? pq[i] : 0.0, rb[i], rf[i], m[i], oth[i],
                          pfb[i], pbf[i]};
              },
              [&](int i, const tl& v, double (&ac)[6]) {
                eq[i] = v.zq;
                eq[i + a.Dpad] = v.zp;
                eq[i + 2 * a.Dpad] = v.zg;
                if (take) sq[i] = v.pq;
                const double r = v.rb + v.rf;
                rho[i] = r;
                const double fpv = fwd ? v.zp : v.oth;
                const double bpv = fwd ? v.oth : v.zp;
                ac[0] += (v.m * fpv) * r;
                ac[1] += (v.m * bpv) * r;
                ac[2] += (v.m * v.pfb) * (v.rb + v.pfb);
                ac[3] += (v.m * bpv) * (v.rb + v.pfb);
                ac[4] += (v.m * fpv) * (v.rf + v.pbf);
                ac[5] += (v.m * v.pbf) * (v.rf + v.pbf);
              });
        }
        block_sum_n<6>(k, acc);
        bool persist = acc[0] > 0 && acc[1] > 0;
        persist &= acc[2] > 0 && acc[3] > 0;
        persist &= acc[4] > 0 && acc[5] > 0;
        __syncthreads();
        if (lane == 0) {
          s.depth++;
          s.lsw = log_sum_exp(s.lsw, s.lsw_sub);
          s.phase = persist ? PH_DOUBLING : PH_TRANS_END;
        }
        break;
      }
      case PH_TRANS_END: {
        // z = z_sample; record; adapt.  One pass: z.q = sample, the draw
        // row (write_draw) and, in an adaptation window, the Welford update
        // with the sample (learn_variance)
        const bool warm = s.iter < a.num_warmup;
        const int mi = warm ? s.iter : s.iter - a.num_warmup;
        const bool draw = (warm ? a.save_warmup != 0 : true) && (mi % a.num_thin) == 0 &&
                          s.rows < a.rows_cap;
        const bool welford = s.adapt_flag && in_adapt_window(k);
        {
          const double n = s.w_n + 1.0;
          const double* sq = k.slot(SQ);
          double* zq = k.slot(ZQ);
          double* wm = k.slot(WM);
          double* wm2 = k.slot(WM2);
          double* row = a.draws + (static_cast<int64_t>(c) * a.rows_cap + s.rows) * a.ncol;
          struct sw { double q, m, m2; };
          vpass<sw>(k,
                    [&](int i) { return sw{sq[i], welford ? wm[i] : 0.0, welford ? wm2[i] : 0.0}; },
                    [&](int i, const sw& v) {
                      zq[i] = v.q;
                      if (draw) row[7 + i] = v.q;
                      if (welford) {
                        const double delta = v.q - v.m;
                        const double mn = v.m + delta / n;
                        wm[i] = mn;
                        wm2[i] = v.m2 + (v.q - mn) * delta;
                      }
                    });
        }
        __syncthreads();
        if (lane == 0) {
          s.accept = s.smp / static_cast<double>(s.n_leapfrog);
          s.V_z = s.V_samp;
          if (draw) {
            if (k.part() == 0) {
              double* row = a.draws + (static_cast<int64_t>(c) * a.rows_cap + s.rows) * a.ncol;
              row[0] = -s.V_samp;
              row[1] = s.accept;
              row[2] = s.eps;
              row[3] = s.depth;
              row[4] = s.n_leapfrog;
              row[5] = s.divergent;
              row[6] = s.T_samp + s.V_samp;
            }
            s.rows++;
          }
        }
        __syncthreads();
        if (s.adapt_flag) {
          if (lane == 0) learn_stepsize(k, s.accept);
          __syncthreads();
          const int upd = learn_variance(k, welford);
          __syncthreads();
          if (upd < 0) {
            finish(k, 70);
          } else if (upd > 0) {
            begin_stepsize_heuristic(k, PH_POST_ADAPT);
          } else if (lane == 0) {
            s.iter++;
            s.phase = PH_TRANS_BEGIN;
          }
        } else if (lane == 0) {
          s.iter++;
          s.phase = PH_TRANS_BEGIN;
        }
        break;
      }
      default:
        finish(k, 70);
        break;
    }
    __syncthreads();
  }
  // a part did not answer an exchange: end the chain (this part failed, or
  // another part's failure is already visible)
  if (P > 1 && lane == 0 && !s.xfail &&
      __hip_atomic_load(a.xfail_g + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    s.xfail = 1;
  __syncthreads();
  if (s.xfail && s.phase != PH_DONE) finish(k, 70);
  __syncthreads();
  if (part == 0) {  // every part holds the same state; one writes it back
    const int* src = reinterpret_cast<const int*>(&s);
    int* dst = reinterpret_cast<int*>(a.states + c);
    for (int i = lane; i < static_cast<int>(sizeof(chain_state) / 4); i += nt) dst[i] = src[i];
  }
}

hipError_t launch_nuts_step(const batch_args& a, hipStream_t s) {
  // one wave per chain for small D, up to 8 waves for large D; parts
  // workgroups per chain
  const int nt = a.D <= 2048 ? 64 : (a.D <= 8192 ? 256 : 512);
  hipLaunchKernelGGL(nuts_step, dim3(a.C * a.parts), dim3(nt), 0, s, a);
  return hipGetLastError();
}

}  // namespace nb
}  // namespace sx
