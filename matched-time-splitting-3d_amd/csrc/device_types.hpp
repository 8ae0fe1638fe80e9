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

// Stage record of a block (kRecLen doubles, one kRecRow group per patched
// row): the patched row's right-hand side of one stage kind written as a
// linear form in the block's five input rows ll-1 .. ll+3 (A: the stage
// input, S: the second input of a Douglas stage) and in the time factors,
//   d_r = sum_t WA[t] A(ll-1+t) + sum_t WS[t] S(ll-1+t)
//         + e_jump K + e_f1 K1 + e_f2 K2,
// then the row's tridiagonal coefficients (a, b, c) after Convert2Tri.  The
// coefficients are derived on the host from the reference-order block
// computation (block_rhs) evaluated on unit inputs.
enum : int {
  kRecWA = 0, kRecWS = 5, kRecK = 10, kRecK1 = 11, kRecK2 = 12, kRecLhs = 13,
  kRecRow = 16, kRecLen = 48,
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
  double ihh[3];          // 1 / (h*h) per axis
  double vdc[3][2][3];    // operator weights of a constant-beta side [axis][0 in,1 out][3]
};

// Device view of the grid: sizes, x pitch (row stride of the field arrays,
// 128 bytes times a whole number (16 doubles, 32 floats) so every x row starts
// on a 128-byte line),
// node classification and the per-axis sin/cos tables.
struct DevGeom {
  int n[3];
  int px;
  const int8_t *ind;
  const double *tab[3];
};

// Device view of one direction's MIB blocks.
// seq: one byte per line, 1 where a patched row of the line is not
// diagonally dominant (|b| < |a| + |c| after Convert2Tri).  Such lines are
// solved with the sequential Thomas algorithm (the reference's TDMA order):
// the partitioned solve is stable only for diagonally dominant lines.  Null
// when no line of the direction needs it.
// cls0: one byte per line, 1 when the line's first node is outside the
// interface.  Along a line the classification flips at every crossing (the
// crossing scan records every sign change), so a kernel recovers the class of
// any row from cls0 and the line's blocks without reading the per-node
// classification.
struct DevMib {
  const int *blk_ptr;
  const BlockHdr *hdr;
  const double *val;
  const double *cr;
  const uint8_t *seq;
  const uint8_t *cls0;
  const double *rec;     // stage records of the launching stage kind (kRecLen per block)
};

// How a line of n nodes is split over the lanes of one lane group.  Rows 0
// and n-1 (boundary values) are kept out of the split; interior rows
// 1..n-2 form p segments of m rows (the last one mlast rows, 2 <= mlast),
// one per lane of a group of L lanes (lanes p..L-1 idle).  In LDS a line
// occupies lds_len double slots: row 0 at slot 0, segment k at
// 1 + k * stride (stride odd, so that the lanes of a group start on distinct
// banks), row n-1 at slot 1 + p * stride.
struct LinePart {
  int n;
  int L;       // lanes per line: 16, 32 or 64
  int mmax;    // register rows per lane (compile-time bound of the kernel)
  int m;       // rows per segment except the last
  int mlast;   // rows of the last segment
  int p;       // segments
  int stride;  // LDS slots per segment
  int magic;   // ceil(2^20 / m): t / m == (t * magic) >> 20 for t < 1024
  int lds_len; // LDS slots per line
};

// Right-hand-side kinds of a line stage.
enum : int {
  kRhsCopy = 0,   // d = A                                  (Douglas x; LOD-IE; TS implicit y/x)
  kRhsDouglas = 1,// d = -dt L_d S + A                      (Douglas y, z)
  kRhsCN = 2,     // d = A + theta dt (L_d A + jumps)        (LOD-CN; TS explicit x, y)
  kRhsTSZ = 3,    // CN form, then + dt/2 f(t^n) + dt/2 f(t^{n+1})   (TS z)
  kFlagLhsJump = 4,   // patched LHS rows carry theta dt vd (jumps) on their RHS
  kFlagSolve = 8,     // tridiagonal solve along the line
  kFlagEpiF = 16,     // epilogue: x + dt f(t)  on every row (LOD last sweep)
  kFlagSeq = 32,      // some lines need the sequential Thomas solve (DevMib::seq)
  kFlagFull = 64,     // every segment but a line's last has exactly mmax rows (set at launch)
  kFlagPre1 = 128,    // Douglas x sweep: the tile holds the stage-1 right-hand side of u^n
};

struct LineArgs {
  DevGeom g;
  DevMib mib;
  LinePart q;          // filled at launch from the line length
  int min_lanes;       // lane-group lower bound for the split (16: the default)
  PhysConst P;
  // Fields, FP64 (double) or FP32 (float) as the instance stores them; the
  // kernel instantiation for that type reads and writes them.
  const void *A;       // stage input
  const void *S;       // second input (Douglas y/z: u^n), may be null
  void *out;           // stage output (may alias A or S)
  double theta_dt;     // implicit coefficient
  double dt;
  double e_out;        // exp(-t) of the output boundary rows
  double e_in;         // exp(-t) of the input boundary rows when analytic
  int in_analytic;     // input rows 0 and n-1 replaced by the analytic boundary value
  double e_jump;       // exp(-t) of the jumps
  double e_f1, c_f1;   // TS z: f(t^n) with c = dt/2;  LOD epilogue: f(t^{n+1}) with c = dt
  double e_f2, c_f2;   // TS z: f(t^{n+1}) with c = dt/2
  // Regular lines' solve tables of the direction (regular_solve.hpp), in the
  // field's type: reg[0] for the guarded solve, reg[1] for the unguarded one
  // (kFlagFull); each holds the outside table, then (reg_sides 2: beta
  // constant inside) the inside one, of reg_side elements each.
  const void *reg[2];
  int reg_side, reg_sides;
  // Solve stages with split set: the stage's tiles split into regular ones
  // (tiles[0], every line regular, at the stage's line split) and the others
  // (tiles[1], at the split of gen_lanes lanes per line: fewer rows per lane,
  // fewer registers, more workgroups per CU for the general solve), entries
  // k * 65536 + bx (see tile_lists), ntiles[c] of each (a list with none is
  // null).  split 0: one launch over every tile.
  int split;
  const int *tiles[2];
  int ntiles[2];
  int gen_lanes;
};

// Douglas stage 1 right-hand side: R = dt L_y u + dt L_z u + u + dt f  (regular
// stencil everywhere; boundary nodes get the boundary value).
struct Rhs1Args {
  DevGeom g;
  PhysConst P;
  const void *u;  // the field: double, or float in the FP32 form
  void *R;        // the stage-1 buffer, of the field's type
  double dt;
  double e;  // exp(-t^{n+1})
};
// Interface diagnostics of one direction at time factor e = exp(-t): per
// crossing, its two fictitious-point errors (el, er) and its jump error (je),
// written at the crossing's index (the arrays already offset by direction).
struct DiagArgs {
  DevGeom g;
  DevMib mib;
  const double *cr;  // jump records, kCrLen per crossing
  int nj, nk;        // lines of the direction: line id k * nj + j
  double e;
  double *el, *er, *je;
};
// Douglas stage 1 at MIB-affected nodes: recompute R with the fictitious-point
// stencils along y and z and the jump terms of all three directions.
struct SpecArgs {
  DevGeom g;
  PhysConst P;
  DevMib mib[3];
  const int *spec;
  int nspec;
  const void *u;  // of the field's type (as Rhs1Args)
  void *R;
  double dt;
  double e;
};
}  // namespace mts
