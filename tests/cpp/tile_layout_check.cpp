// Host check of the staged-tile LDS layout (stan_amd/csrc/tile_layout.hpp)
// against the operand layouts of v_mfma_f32_32x32x16_bf16:
//  - staging (batched GLM pass): LDS chunk p of a tile holds global chunk
//    swz(r, p % CPR) of row r = p / CPR, for every Kp = 32 .. 256;
//  - element writes: elem_offset(r, c);
//  - chunk copies from a split copy (hierarchical split-product pass, every
//    K = 0 .. 31): split_chunk_offset puts chunk e of a tile whose rows hold
//    3 terms x NCH 16-byte chunks at term (e % 3 NCH) / NCH, row e / 3 NCH,
//    columns 8 (e % NCH) .. + 7, i.e. at elem_offset of each of its eight
//    elements, inside the three term tiles, every element once;
//  - forward operand: lane (i, h) reads, for k-step s, the 8 columns
//    16 s + 8 h .. + 7 of row i;
//  - adjoint operand: two transposing reads ds_read_b64_tr_b16 per column
//    block kb, half st: element j of lane l is row 16 st + 8 (j / 4) +
//    4 (l / 32) + j % 4, column 32 kb + l % 32.  Within each 16-lane group
//    the instruction hands lane i element e of the 8 bytes lane 4 e + i / 4
//    of the group read, at position i % 4.
// Prints "ok" or the first mismatch.
#include <cstdio>
#include <utility>
#include <vector>

#include "../../stan_amd/csrc/tile_layout.hpp"

using sx::elem_offset;
using sx::split_chunk_offset;
using sx::split_chunk_row;
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

template <int K>
int check_split_copy() {
  constexpr int NCH = (K + 8) / 8, ROWB = 64, TB = 32 * ROWB, TC = 32 * 3 * NCH;
  std::vector<int> hits(3 * TB / 2, 0);
  for (int e = 0; e < TC; ++e) {
    const int o = split_chunk_offset<ROWB, 4, NCH>(e, TB);
    const int row = e / (3 * NCH), w = e % (3 * NCH), term = w / NCH;
    if (split_chunk_row<NCH>(e) != row || o < 0 || o + 16 > 3 * TB || o % 16 != 0) {
      std::printf("K=%d chunk %d -> offset %d row %d\n", K, e, o, split_chunk_row<NCH>(e));
      return 1;
    }
    for (int k = 0; k < 8; ++k) {
      const int col = 8 * (w % NCH) + k;
      if (o + 2 * k != term * TB + elem_offset<ROWB, 4>(row, col)) {
        std::printf("K=%d chunk %d element %d not at (%d, %d) of term %d\n", K, e, k, row, col,
                    term);
        return 1;
      }
      hits[(o + 2 * k) / 2]++;
    }
  }
  for (int t = 0; t < 3; ++t)
    for (int r = 0; r < 32; ++r)
      for (int c = 0; c < 32; ++c) {
        const int n = hits[(t * TB + elem_offset<ROWB, 4>(r, c)) / 2];
        if (n != (c < 8 * NCH ? 1 : 0)) {
          std::printf("K=%d term %d (%d, %d) written %d times\n", K, t, r, c, n);
          return 1;
        }
      }
  return 0;
}

template <int... Ks>
int check_split_copies(std::integer_sequence<int, Ks...>) {
  return (check_split_copy<Ks>() + ...);
}

int main() {
  int bad = check_glm<1>() + check_glm<2>() + check_glm<3>() + check_glm<4>() + check_glm<5>() +
            check_glm<6>() + check_glm<7>() + check_glm<8>();
  bad += check_split_copies(std::make_integer_sequence<int, 32>{});
  std::printf(bad ? "mismatch\n" : "ok\n");
  return bad ? 1 : 0;
}
