// Host check of the staged-tile LDS layout (stan_amd/csrc/tile_layout.hpp)
// against the operand layouts of v_mfma_f32_32x32x16_bf16:
//  - staging (batched GLM pass): LDS chunk p of a tile holds global chunk
//    swz(r, p % CPR) of row r = p / CPR, for every Kp = 32 .. 256;
//  - element writes (hierarchical split-product pass): elem_offset(r, c);
//  - forward operand: lane (i, h) reads, for k-step s, the 8 columns
//    16 s + 8 h .. + 7 of row i;
//  - adjoint operand: two transposing reads ds_read_b64_tr_b16 per column
//    block kb, half st: element j of lane l is row 16 st + 8 (j / 4) +
//    4 (l / 32) + j % 4, column 32 kb + l % 32.  Within each 16-lane group
//    the instruction hands lane i element e of the 8 bytes lane 4 e + i / 4
//    of the group read, at position i % 4.
// Prints "ok" or the first mismatch.
#include <cstdio>
#include <vector>

#include "../../stan_amd/csrc/tile_layout.hpp"

using sx::elem_offset;
using sx::swz;
using sx::tile_addr;
using sx::xor_block;

struct cell { int row, col; };

template <int KQ>
int check_glm() {
  constexpr int Kp = 32 * KQ, ROWB = 2 * Kp, CPR = Kp / 8, P = xor_block<CPR>();
  // LDS byte -> (row, column) of the staged tile, from the staging rule
  std::vector<cell> lds(32 * ROWB / 2, cell{-1, -1});
  for (int p = 0; p < 32 * CPR; ++p) {
    const int r = p / CPR, c = swz<P>(r, p % CPR);
    for (int e = 0; e < 8; ++e) lds[(16 * p) / 2 + e] = cell{r, 8 * c + e};
  }
  // element writes give the same map
  for (int r = 0; r < 32; ++r)
    for (int c = 0; c < Kp; ++c) {
      const int o = elem_offset<ROWB, P>(r, c);
      const cell x = lds[o / 2];
      if (x.row != r || x.col != c) {
        std::printf("Kp=%d elem_offset(%d, %d) -> (%d, %d)\n", Kp, r, c, x.row, x.col);
        return 1;
      }
    }
  for (int lane = 0; lane < 64; ++lane) {
    const tile_addr<ROWB, P> ad(lane);
    const int li = lane & 31, h = lane >> 5;
    for (int s = 0; s < 2 * KQ; ++s)
      for (int j = 0; j < 8; ++j) {
        const cell x = lds[ad.fwd(s) / 2 + j];
        if (x.row != li || x.col != 16 * s + 8 * h + j) {
          std::printf("Kp=%d fwd lane %d s %d j %d -> (%d, %d)\n", Kp, lane, s, j, x.row, x.col);
          return 1;
        }
      }
  }
  for (int st = 0; st < 2; ++st)
    for (int kb = 0; kb < KQ; ++kb)
      for (int q = 0; q < 2; ++q)
        for (int lane = 0; lane < 64; ++lane) {
          const int g = lane >> 4, i = lane & 15;
          for (int e = 0; e < 4; ++e) {
            const tile_addr<ROWB, P> src(16 * g + 4 * e + i / 4);
            const int o = src.tr(st, q, kb);
            if (o < 0 || o + 8 > 32 * ROWB || o % 8 != 0) {
              std::printf("Kp=%d tr address %d out of the tile\n", Kp, o);
              return 1;
            }
            const cell x = lds[o / 2 + i % 4];
            const int row = 16 * st + 8 * q + 4 * (lane / 32) + e, col = 32 * kb + lane % 32;
            if (x.row != row || x.col != col) {
              std::printf("Kp=%d tr st %d q %d kb %d lane %d e %d -> (%d, %d), want (%d, %d)\n",
                          Kp, st, q, kb, lane, e, x.row, x.col, row, col);
              return 1;
            }
          }
        }
  return 0;
}

int main() {
  int bad = check_glm<1>() + check_glm<2>() + check_glm<3>() + check_glm<4>() + check_glm<5>() +
            check_glm<6>() + check_glm<7>() + check_glm<8>();
  std::printf(bad ? "mismatch\n" : "ok\n");
  return bad ? 1 : 0;
}
