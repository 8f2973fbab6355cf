// Internal definitions shared by the HIP kernels and the C-ABI layer.
#ifndef STANX_INTERNAL_HPP
#define STANX_INTERNAL_HPP

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/stanx.h"

namespace sx {

void set_error(const std::string& msg);

#define SX_HIP(call)                                                     \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      ::sx::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
      return SX_EDEVICE;                                                 \
    }                                                                    \
  } while (0)

// Geometry of the lane-slot tiled X layout used by the bulk passes.
// A wave-tile holds R = 64 / L rows; each row is split across L lanes and
// each lane holds S = ceil(K / L) columns ("slots").  Element (row, col)
// of the tile t lives at  t*S*64 + (col / L)*64 + (row % R)*L + col % L,
// so slot j of all 64 lanes is one contiguous 512-byte (FP64) segment.
struct tile_geom {
  int L = 1;       // lanes per row (power of two, <= 64)
  int lgL = 0;     // log2(L)
  int R = 64;      // rows per tile
  int S = 0;       // slots per lane
  int smax = 8;    // compiled register capacity (>= S)
  int prefetch = 0;  // 1: next tile's loads issued before this tile's math
  int64_t T = 0;   // number of tiles
};

tile_geom make_geom(int64_t N, int64_t K, int L, int smax);

// Arguments of the FP64 bulk pass (K1 logistic, K2 normal).
struct pass_args {
  const double* X;      // tiled
  const void* y;        // int8 (logistic) or double (normal)
  const double* theta;  // D doubles per chain on the device
  double* partials;     // blocks x E
  double* out;          // lp, grad[D], status
  int64_t N;
  int64_t T;
  int K, L, lgL, S, D, E;
  int propto, jacobian;
  int chains = 1;              // batched evaluation: chains in the grid
  int64_t theta_stride = 0;    // doubles between chains' theta
  int64_t out_stride = 0;      // doubles between chains' outputs (>= D+2)
  double alpha_sd, beta_sd, sigma_sd, sigma_mu;
  // normalising constants -log(sd) - log(2 pi)/2 of the normal priors
  double alpha_const, beta_const, sigma_const;
};

// Bulk X pass; with reduce, also the cross-block reduction + finalize
// launch that writes out[].
hipError_t launch_pass_f64(int family, const tile_geom& g, const pass_args& a,
                           int blocks, int threads, hipStream_t s, bool reduce = true);
hipError_t launch_tile_x_f64(const double* Xrm, double* Xt, int64_t N,
                             int64_t K, const tile_geom& g, hipStream_t s);

// Arguments of the batched BF16 hi+lo MFMA pass.  BF16 arrays are stored
// as uint16_t on the host side.
struct mx_args {
  const uint16_t *Xh, *Xl;  // k-packed X, Np x Kp
  const void* y;
  const float* yf;          // Np: sign (logistic) or value (normal), 0 past N
  const uint16_t *Bh, *Bl, *Bm;  // Cp x Kp chain coefficients (three terms)
  const float* alpha;       // Cp intercepts
  float* gpart;             // splits x Cp x Kp
  double* spart;            // splits x Cp x 3
  uint16_t *Rh, *Rl;        // Kp > 256: BF16 hi/lo residuals, Np x Cp each
  int64_t N, n_tiles, tiles_per_split;
  int Kp, Cp, splits;
  int variant;  // Kp <= 256: 0 one wave per 32-chain block, 1 a k-split pair
};

hipError_t launch_split_x(const double* Xrm, uint16_t* Xh, uint16_t* Xl, int64_t N, int64_t K,
                          int64_t Np, int64_t Kp, hipStream_t s);
hipError_t launch_prep_y(const void* y, float* yf, int family, int64_t N, int64_t Np,
                         hipStream_t s);
// prep (theta -> bf16 coefficients), MFMA pass, FP64 reduction + finalize
// into pa.out (stride pa.out_stride) for C chains
hipError_t launch_mfma_pass(int family, const double* theta, int64_t D, int C,
                            const mx_args& a, const pass_args& pa, hipStream_t s);

// Hierarchical logistic model: group-aligned row chunks of at most
// `chunk` rows (kHierChunk by default; the autotuner picks one of
// 128 .. 1024 per data shape); partials per chunk and per group.
constexpr int kHierChunk = 256;
constexpr int kHierChunkMin = 128;
struct hier_args {
  const double* X;          // N x K row-major (FP64 pass)
  const float* X32;         // N x K row-major (FP32 batched pass)
  const uint16_t* XS;       // (N + 32) x 3 x 8 hier_x6_chunks(K) BF16 (x6 pass)
  uint16_t* wsplit;         // J x chains x 3 x 32 BF16: split coefficients (x6 pass)
  const int8_t* y;
  const double* theta;      // chains x theta_stride
  double* out;              // chains x out_stride: lp, grad[D], status
  double* cpart;            // n_chunks x chains x E
  double* gpart;            // chains x J x G
  const int* chunk_group;   // n_chunks
  const int64_t* chunk_row0;    // n_chunks + 1
  const int64_t* group_chunk0;  // J + 1
  int64_t n_chunks, theta_stride, out_stride;
  int J, K, D, E, G, chains, propto, jacobian;
  int x6;  // FP32 batches: products on BF16 MFMA in three-term splits (K <= 31)
  double alpha_sd, beta_sd, tau_sd, alpha_const, beta_const, tau_const;
};
hipError_t launch_hier(const hier_args& a, bool fp32, hipStream_t s);
// 16-byte chunks of BF16 per row and term of the x6 pass's split copy of
// X: columns 0 .. K (x' and the sign column), zero-padded to a multiple of 8
constexpr int hier_x6_chunks(int K) { return (K + 8) / 8; }

// Device workspace of batched MFMA evaluations for up to cap chains.
struct mx_workspace {
  int cap = 0, Cp = 0, splits = 0;
  uint16_t *Bh = nullptr, *Bl = nullptr, *Bm = nullptr;
  uint16_t *Rh = nullptr, *Rl = nullptr;  // Kp > 256 only
  float* alpha = nullptr;
  float* gpart = nullptr;
  double* spart = nullptr;
};

// Partial sums of hierarchical evaluations for up to cap chains.
struct hier_workspace {
  int64_t cap = 0;
  double* cpart = nullptr;  // n_chunks x chains x (K + 3)
  double* gpart = nullptr;  // chains x J x (2K + 5)
  uint16_t* wsplit = nullptr;  // J x chains x 3 x 32 (FP32 contexts, K <= 31)
};

// Each user of batched evaluations (the context's own batch calls, every
// batched sampler) owns its workspace: a captured graph keeps the
// addresses it was captured with, and samplers may run concurrently.
void free_workspace(mx_workspace& w);
void free_workspace(hier_workspace& w);

}  // namespace sx

struct sx_glm;
namespace sx {
int mx_enqueue(sx_glm* m, mx_workspace& w, const double* d_theta, int C, double* d_out,
               int propto, int jacobian, hipStream_t stream);
int hier_prepare(sx_glm* m, hier_workspace& w, int C);
int mx_autotune(sx_glm* m);
int plain_enqueue(sx_glm* m, const double* d_theta, double* d_out, double* partials, int propto,
                  int jacobian, hipStream_t stream, bool reduce);
// the per-device scratch array of the allocation-free entry point: its
// address on `device`, looked up once (called at context creation)
int noalloc_setup(int device);
int hier_enqueue(sx_glm* m, hier_workspace& w, const double* d_theta, int C, double* d_out,
                 int propto, int jacobian, bool fp32, hipStream_t stream);
}

struct sx_glm {
  int family = 0, precision = 0, device = 0;
  int64_t N = 0, K = 0, J = 0, D = 0;
  sx_prior prior{};
  hipStream_t stream = nullptr;
  sx::tile_geom geom;
  double* d_X = nullptr;
  void* d_y = nullptr;
  double* d_theta = nullptr;
  double* d_partials = nullptr;
  double* d_out = nullptr;
  double* h_theta = nullptr;  // pinned
  double* h_out = nullptr;    // pinned
  int blocks = 0, threads = 256, max_blocks = 0;
  // BF16 hi+lo layouts for batched MFMA evaluation (precision SX_BF16X2)
  int64_t Np = 0, Kp = 0;
  uint16_t *d_Xh = nullptr, *d_Xl = nullptr;  // k-packed BF16 hi and lo
  float* d_yf = nullptr;                       // y per padded row
  sx::mx_workspace mx;
  // launch shape of the MFMA pass (fixed per context; chosen by the
  // autotuner): kernel variant and target tiles of 32 rows per row split
  int mx_variant = 0, mx_tps = 8;
  double* d_btheta = nullptr;  // batched theta / out staging for the C-ABI batch call
  double* d_bout = nullptr;
  int64_t bcap = 0;
  // hierarchical model
  float* d_X32 = nullptr;
  uint16_t* d_XS = nullptr;  // split copy of the sign-folded X (FP32, K <= 31)
  int* d_chunk_group = nullptr;
  int64_t *d_chunk_row0 = nullptr, *d_group_chunk0 = nullptr;
  int64_t n_chunks = 0;
  int hier_chunk = 0;  // rows per chunk at most
  int hier_x6 = 0;     // FP32 batches: products in BF16 three-term splits
  sx::hier_workspace hw;
  std::string tuning_log;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

#endif
