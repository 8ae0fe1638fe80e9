// Plain-old-data shared by the host table builder and the HIP kernels.
#pragma once

#include <cstdint>

namespace mts {

// Sparse MIB table, per line direction d (0 = x, 1 = y, 2 = z).
//
// A "block" is what one pass of the reference's per-line MIB loop handles:
// an irregular crossing (rows ll, ll+1) or a corner pair (rows ll..ll+2).
// Block record layout in blk_val (doubles):
//   irregular (kBlkIrrLen):  vd[2][3] | wl[6] | wr[6] | lhs[2][3] | mult[2]
//   corner    (kBlkCorLen):  vd[3][3] | Lwl[9] | Lwr[9] | Rwl[9] | Rwr[9] | lhs[3][3] | mult[6]
// vd:   operator weights beta(x -+ h/2)/h^2 of each patched row (Operator_weights_*).
// w*:   fictitious-point weights (Appendix-A layout of the reference's Wei).
// lhs:  (a, b, c) of each patched row after Convert2Tri_irr / Convert2Tri_cor.
// mult: the row-elimination multipliers of Convert2Tri_*, replayed on the
//       right-hand side every step in the reference's order:
//         irregular: d1 -= m1 d0; d0 -= m2 d1
//         corner:    d1 -= m10 d0; d2 -= m20 d0; d0 -= m02 d2; d1 -= m12 d2;
//                    d2 -= m21 d1; d0 -= m01 d1
enum : int {
  kBlkIrr = 1,
  kBlkCor = 2,
  kBlkIrrLen = 26,
  kBlkCorLen = 60,
  // offsets inside an irregular record
  kIrrVd = 0, kIrrWl = 6, kIrrWr = 12, kIrrLhs = 18, kIrrMul = 24,
  // offsets inside a corner record
  kCorVd = 0, kCorLwl = 9, kCorLwr = 18, kCorRwl = 27, kCorRwr = 36, kCorLhs = 45, kCorMul = 54,
  // per-crossing jump record (doubles): sin x,y,z | cos x,y,z | beta-, beta+ | normal[3] | P^-1[d][0]
  kCrLen = 12,
  // special-node record for the Douglas stage-1 right-hand side (ints):
  // ix, iy, iz, blk[3], row[3], pad
  kSpecLen = 10,
};

// Block header (4 ints): type, ll, first crossing index, offset into blk_val.
struct BlockHdr {
  int type, ll, c0, off;
};

// Axis tables, each of length n[a]: sin/cos of the nodes and sin/cos of
// node -/+ h/2 (the half points the operator weights evaluate beta at).
enum : int { kTabSin = 0, kTabCos, kTabSinM, kTabSinP, kTabCosM, kTabCosP, kTabCount };

// Physics constants for one instance (equation: e^{-t} sin sin sin inside,
// e^{-t} cos cos cos outside; beta kind 0 constant, 1 c(1 + 0.5 sin x cos y) inside).
struct PhysConst {
  int beta_kind;
  double b_in, b_out;     // beta- (or its amplitude c) and beta+
  double f_in, f_out;     // 3*beta - 1 for the constant sides
  double hh[3];           // h*h per axis
  double vdc[3][2][3];    // operator weights of a constant-beta side [axis][0 in,1 out][3]
};

}  // namespace mts
