// Right-hand-side-only line solve for regular lines.
//
// A line without MIB blocks whose side has a constant diffusion coefficient
// has the same tridiagonal matrix in every row (a = -theta dt beta/h^2,
// b = 1 + 2 theta dt beta/h^2, c = a) apart from its ends, whose couplings
// to the boundary values are moved to the right-hand side.  Every such line
// of a direction and side therefore has the same matrix, and the matrix does
// not change from step to step.  The partitioned solve of the line kernels
// (each lane eliminates its segment of rows keeping its first unknown as a
// parameter, the first unknowns are folded out, the one-unknown-per-lane
// system is solved by cyclic reduction over the lane group, each lane
// back-substitutes) splits into a part that depends on the matrix only and a
// part that is linear in the right-hand side.  The matrix part is computed
// once per instance on the host (regular_factors) and staged in LDS; the
// kernels then run only the right-hand-side part (reg_forward, reg_fold,
// reg_back), a few fused multiply-adds per row.  Both parts follow the
// operation order of the kernels' general partitioned solve.
#pragma once

#include <cstdint>
#include <vector>

#include "device_types.hpp"

#if defined(__HIPCC__)
#define MTS_RS_HD __host__ __device__ __forceinline__
#else
#define MTS_RS_HD inline
#endif

namespace mts {

// Per-row factors of one segment variant (kRegRow doubles, 16-byte aligned):
//   r     reciprocal pivot of the downward elimination (row >= 1)
//   a     the row's sub-diagonal coupling (0 on padding rows)
//   cf    the row's normalized super-diagonal after the downward pass
//         (used by the upward pass on the right-hand side)
//   af    coefficient of the segment's first unknown after the upward pass
//   cfin  coefficient of the segment's last unknown after the upward pass
enum : int { kRegR = 0, kRegA = 1, kRegCf = 2, kRegAf = 3, kRegCfin = 4, kRegRow = 6 };
// Per-lane scalars (kRegScal doubles):
//   cD    the coupling of the segment's first row to its second (0 when the
//         segment has two rows: the first row then stays as it is)
//   i1    reciprocal of the first row's reduced pivot
//   A1, C1 the first row's couplings to the previous segment's last unknown
//         and to this segment's last unknown
//   p1, p2 the fold's weights of this segment's and the next segment's
//         first-row right-hand sides
//   k1[l], k2[l] cyclic-reduction weights of level l (shift 1, 2, 4, ...)
//   q     reciprocal of the lane's final reduced pivot
enum : int {
  kRegCD = 0, kRegI1 = 1, kRegA1 = 2, kRegC1 = 3, kRegP1 = 4, kRegP2 = 5, kRegK = 6,
  kRegMaxLevels = 6, kRegQ = kRegK + 2 * kRegMaxLevels, kRegScal = kRegQ + 3,
};

// Table of one direction and side: the common segment (every lane but the
// last) and the last segment's row factors, then every lane's scalars.  The
// layout keeps the reads of one instruction on different LDS banks: the
// lanes of a group read one row of the common segment (a broadcast) and the
// last lane the same row of its own segment, kRegPad doubles off the bank of
// the first; each lane reads the same scalar of its own block, an odd number
// of doubles (kRegScal) apart.
constexpr int kRegPad = 2;
MTS_RS_HD int reg_side_len(int L, int mmax) {  // (even: L is a power of two >= 16)
  return 2 * mmax * kRegRow + 2 * kRegPad + L * kRegScal;
}
MTS_RS_HD int reg_rows_off(int variant, int mmax) { return variant * (mmax * kRegRow + kRegPad); }
MTS_RS_HD int reg_scal_off(int lane, int mmax) {
  return 2 * mmax * kRegRow + 2 * kRegPad + lane * kRegScal;
}

// Downward and upward elimination of the lane's right-hand side d[0..m)
// (FULL: m == MMAX, padding rows hold d = 0).  Returns the segment's reduced
// right-hand sides D1 (first row) and D2 (last row).
template <int MMAX, bool FULL, typename R>
MTS_RS_HD void reg_forward(const R *rows, const R *sc, R (&d)[MMAX], int m_lane, R &D1, R &D2) {
  const int m = FULL ? MMAX : m_lane;
  D1 = R(0);
  D2 = R(0);
  if (m < 2) return;
  d[1] = d[1] * rows[1 * kRegRow + kRegR];
  D2 = d[1];
#pragma unroll
  for (int i = 2; i < MMAX; i++) {
    if (FULL || i < m) {
      const R *w = rows + i * kRegRow;
      d[i] = fma(-w[kRegA], d[i - 1], d[i]) * w[kRegR];
      D2 = d[i];
    }
  }
#pragma unroll
  for (int i = MMAX - 3; i >= 1; i--) {
    if (FULL || i <= m - 3) d[i] = fma(-rows[i * kRegRow + kRegCf], d[i + 1], d[i]);
  }
  D1 = fma(-sc[kRegCD], d[1], d[0]);
}

// The fold: the lane's equation in the segment-end unknowns (right-hand side
// only); D1n is the next lane's D1 (0 past the group).
template <typename R>
MTS_RS_HD R reg_fold(const R *sc, int m_lane, R D1, R D2, R D1n) {
  if (m_lane < 2) return R(0);
  return fma(-sc[kRegP2], D1n, fma(-sc[kRegP1], D1, D2));
}

// One level of cyclic reduction: dm and dp are the right-hand sides of the
// lanes `shift` below and above (0 outside the group).
template <typename R>
MTS_RS_HD R reg_level(const R *sc, int level, R ed, R dm, R dp) {
  return fma(-dp, sc[kRegK + 2 * level + 1], fma(-dm, sc[kRegK + 2 * level], ed));
}

// Back-substitution: xe = ed q is the segment's last unknown, xprev the
// previous segment's; on return d[0..m) holds the lane's solution.
template <int MMAX, bool FULL, typename R>
MTS_RS_HD void reg_back(const R *rows, const R *sc, R (&d)[MMAX], int m_lane, R D1, R xe,
                        R xprev) {
  const int m = FULL ? MMAX : m_lane;
  const R xs = fma(-sc[kRegC1], xe, fma(-sc[kRegA1], xprev, D1)) * sc[kRegI1];
  const bool act = m >= 2;
#pragma unroll
  for (int i = 0; i < MMAX; i++) {
    const R *w = rows + i * kRegRow;
    const R mid = fma(-w[kRegCfin], xe, fma(-w[kRegAf], xs, d[i]));
    const R v = i == 0 ? xs : (i == m - 1 ? xe : mid);
    d[i] = act && i < m ? v : d[i];
  }
}

// Host: the tiles of one line stage along direction d split by class.  A
// tile is kernel workgroup (bx, k): lines j = bx T .. bx T + T - 1 (T = 256 /
// L) of outer index k; it is regular when every interior line in it is
// regular (no MIB block; outside, or inside with a constant coefficient).
// Entries are k * 65536 + bx, in launch order (k outer, bx inner); interior k
// only.  blk_ptr and cls0 are the direction's table (n_lines + 1 and n_lines
// entries, line id k * nj + j).
void tile_lists(const int *blk_ptr, const uint8_t *cls0, bool beta_const, int nj, int nk, int L,
                std::vector<int> &regular, std::vector<int> &general);
// The general tiles re-cut for a wider lane group: each tile of T_from
// lines becomes T_from / T_to tiles of T_to lines (T_to divides T_from),
// leaving out those past the last line nj - 1.
void recut_tiles(const std::vector<int> &from, int t_from, int t_to, int nj, std::vector<int> &out);
// Tiles that share one 128-byte line: y/z tiles are T consecutive x, so a
// tile of T elements of elem_bytes narrower than 128 bytes is one of
// 128 / (T elem_bytes) per line (1: whole lines).
MTS_RS_HD constexpr int line_group(int dir, int T, int elem_bytes) {
  return (dir != 0 && T * elem_bytes < 128) ? 128 / (T * elem_bytes) : 1;
}
// A tile list reordered so that the tiles of one line group (same k, same
// bx / group) are 8 entries apart: the workgroups of a launch go to the 8
// XCDs in turn, so those tiles run on one XCD, dispatched one after the
// other, and the line is read from HBM once (the others find it in that
// XCD's L2).  Groups of eight full line groups first, then the groups with
// tiles in the other list, in order.  A permutation: every tile stays.
void xcd_group_tiles(std::vector<int> &list, int group);
// Both lists of one solve stage along direction dir of n nodes: the regular
// tiles at the split of min_lanes lanes, the general ones at the split of
// max(gen_lanes, min_lanes) lanes, each grouped by line for elements of
// elem_bytes (xcd_group_tiles).  False when n has no split.
bool stage_tiles(const int *blk_ptr, const uint8_t *cls0, bool beta_const, int dir, int n, int nj,
                 int nk, int min_lanes, int gen_lanes, int elem_bytes, std::vector<int> &regular,
                 std::vector<int> &general);
MTS_RS_HD int tile_bx(int e) { return e & 0xffff; }
MTS_RS_HD int tile_k(int e) { return e >> 16; }

// Host: the matrix part of the partitioned solve for the lines of one split
// q whose rows are all (a, b, c), the first row's a and the last row's c
// moved to the right-hand side.  full: the kernel's unguarded variant (every
// lane eliminates mmax rows, the rows past a lane's segment are identity
// rows).  Appends reg_side_len(q.L, mmax) doubles to out.  Returns false
// when a pivot is zero.
bool regular_factors(const LinePart &q, int mmax, bool full, double a, double b, double c,
                     std::vector<double> &out);

}  // namespace mts
