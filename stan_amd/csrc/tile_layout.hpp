// Address arithmetic of the staged-tile LDS layout shared by the BF16
// matrix-core passes: plain integer code, compiled for the device by the
// kernels and for the host by the layout checks.
#ifndef STANX_TILE_LAYOUT_HPP
#define STANX_TILE_LAYOUT_HPP

#if defined(__HIPCC__) || defined(__HIP__)
#define SX_HD __host__ __device__ __forceinline__
#else
#define SX_HD inline
#endif

namespace sx {

// Staged tile layout in LDS: row r of a tile occupies ROWB bytes; its
// 16-byte chunk c (8 consecutive columns) sits at chunk position
// c ^ f(r), f(r) = 2 (r & (P/2 - 1)), P the largest power of two <= 16
// dividing the chunks per row.  The XOR spreads both read patterns over the
// banks: the forward's 16-byte reads (8 consecutive rows, one chunk each)
// and the transposing reads (4 consecutive rows x 2 adjacent chunks per 16
// lanes).  Because f(r) only changes bits 1..3 of the chunk index and
// depends on r mod 8, every read address is a per-lane base, one XOR with
// a per-lane value and an immediate.
template <int ROWB, int P>
struct tile_addr {
  static_assert(P == 4 || P == 8 || P == 16, "XOR block of 4, 8 or 16 chunks");
  static constexpr int GM = P / 2 - 1;
  int fb, fx, tb, tx;
  SX_HD explicit tile_addr(int lane) {
    const int li = lane & 31, h = lane >> 5;
    fb = li * ROWB + 16 * h;
    fx = 32 * (li & GM);
    // transposing reads: within each group of 16 lanes, lane i reads 8
    // bytes of row 4 hh + i / 4 (plus 16 st + 8 q), columns c0 + 16 ch +
    // 4 (i % 4) .. + 3
    const int i = lane & 15, g = lane >> 4, hh = g >> 1, ch = g & 1;
    const int rr = 4 * hh + (i >> 2);
    const int f = 2 * (rr & GM);
    const int cl = 2 * ch + ((i & 3) >> 1);
    tb = rr * ROWB + 16 * (cl ^ (f & 3)) + 8 * (i & 1);
    tx = 16 * (f & ~3);
  }
  // byte offset of the forward operand of k-step s (columns 16 s + 8 h ..)
  SX_HD int fwd(int s) const {
    return fb + ((32 * (s % (P / 2))) ^ fx) + 16 * P * (s / (P / 2));
  }
  // byte offset of the transposing read q of half st, column block kb
  SX_HD int tr(int st, int q, int kb) const {
    return tb + (16 * st + 8 * q) * ROWB + 16 * P * (kb / (P / 4)) +
           ((64 * (kb % (P / 4))) ^ tx);
  }
};

// chunk position of global chunk c of tile row r (and the reverse)
template <int P>
SX_HD int swz(int r, int c) { return c ^ (2 * (r & (P / 2 - 1))); }

template <int CPR>
constexpr int xor_block() { return CPR % 16 == 0 ? 16 : CPR % 8 == 0 ? 8 : 4; }

// byte offset of the BF16 element (r, c) of a tile (global chunk c / 8 of
// row r at its permuted position)
template <int ROWB, int P>
SX_HD int elem_offset(int r, int c) { return r * ROWB + 16 * swz<P>(r, c >> 3) + 2 * (c & 7); }

// Staging from a split copy whose rows hold 3 terms x NCH 16-byte chunks
// (term-major, columns 0 .. 8 NCH - 1 of each term): chunk e of a 32-row
// tile lies in row e / (3 NCH) of term (e % (3 NCH)) / NCH, at column
// chunk e % NCH.  Byte offset of its destination within the three term
// tiles (TB bytes apart), and its tile row.
template <int ROWB, int P, int NCH>
SX_HD int split_chunk_offset(int e, int TB) {
  const int r = e / (3 * NCH), w = e % (3 * NCH);
  return (w / NCH) * TB + elem_offset<ROWB, P>(r, 8 * (w % NCH));
}
template <int NCH>
SX_HD int split_chunk_row(int e) { return e / (3 * NCH); }

}  // namespace sx

#endif
