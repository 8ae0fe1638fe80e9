// gfx950 line-stage kernels (see kernels.hip for the overview).  This file
// is compiled once per line direction and lane-group size
// (-DMTS_LINE_DIR=0|1|2 -DMTS_LINE_L=16|32|64), so the many template
// instances build in parallel; each object defines one
// launch_line_dl_<dir>_<L> entry used by launch_line.
#include <hip/hip_runtime.h>

#include "device_core.hpp"
#include "kernels.hpp"
#include "regular_solve.hpp"

#ifndef MTS_LINE_DIR
#error "compile with -DMTS_LINE_DIR=0|1|2 -DMTS_LINE_L=16|32|64"
#endif

namespace mts {

namespace {

constexpr int kBlock = kBlockThreads;

// ------------------------------------------------- group line solver
// Lanes of a group of L lanes (16, 32 or 64; a DPP row when L = 16) hold the
// segments of one line.  Shifts within the group return id where the source
// lane lies outside it.
template <int L>
__device__ __forceinline__ double grp_up(double v, int s, int kseg, double id) {
  if (L == 16) {
    // row_shr:s -- lane i takes lane i - s of its 16-lane row
    const int ctl = 0x110 | s;
    int lo = 0, hi = 0;
    switch (ctl) {
#define MTS_DPP(c)                                                                        \
  case c:                                                                                 \
    lo = __builtin_amdgcn_update_dpp(__double2loint(id), __double2loint(v), c, 0xf, 0xf, false); \
    hi = __builtin_amdgcn_update_dpp(__double2hiint(id), __double2hiint(v), c, 0xf, 0xf, false); \
    break;
      MTS_DPP(0x111) MTS_DPP(0x112) MTS_DPP(0x114) MTS_DPP(0x118)
      default: break;
    }
    return __hiloint2double(hi, lo);
  } else {
    const double w = __shfl_up(v, s, L);
    return kseg < s ? id : w;
  }
}
template <int L>
__device__ __forceinline__ double grp_down(double v, int s, int kseg, double id) {
  if (L == 16) {
    // row_shl:s -- lane i takes lane i + s of its 16-lane row
    const int ctl = 0x100 | s;
    int lo = 0, hi = 0;
    switch (ctl) {
      MTS_DPP(0x101) MTS_DPP(0x102) MTS_DPP(0x104) MTS_DPP(0x108)
      default: break;
    }
    return __hiloint2double(hi, lo);
  } else {
    const double w = __shfl_down(v, s, L);
    return kseg + s >= L ? id : w;
  }
}
template <int L>
__device__ __forceinline__ float grp_up(float v, int s, int kseg, float id) {
  if (L == 16) {
    const int ctl = 0x110 | s;
    int w = 0;
    switch (ctl) {
#define MTS_DPP1(c)                                                                          \
  case c:                                                                                    \
    w = __builtin_amdgcn_update_dpp(__float_as_int(id), __float_as_int(v), c, 0xf, 0xf, false); \
    break;
      MTS_DPP1(0x111) MTS_DPP1(0x112) MTS_DPP1(0x114) MTS_DPP1(0x118)
      default: break;
    }
    return __int_as_float(w);
  } else {
    const float w = __shfl_up(v, s, L);
    return kseg < s ? id : w;
  }
}
template <int L>
__device__ __forceinline__ float grp_down(float v, int s, int kseg, float id) {
  if (L == 16) {
    const int ctl = 0x100 | s;
    int w = 0;
    switch (ctl) {
      MTS_DPP1(0x101) MTS_DPP1(0x102) MTS_DPP1(0x104) MTS_DPP1(0x108)
      default: break;
    }
    return __int_as_float(w);
  } else {
    const float w = __shfl_down(v, s, L);
    return kseg + s >= L ? id : w;
  }
}
#undef MTS_DPP1
#undef MTS_DPP

// 1 / x: hardware reciprocal refined by two Newton steps (within an ulp or
// so of the correctly rounded quotient).
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// 1 / x in FP32: hardware reciprocal refined by one Newton step.
__device__ __forceinline__ float rcp_nr(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r, 1.0f);
  return fmaf(r, e, r);
}
__device__ __forceinline__ double fmar(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float fmar(float a, float b, float c) { return fmaf(a, b, c); }

// Partitioned solve of the group's line: lane k holds rows [s_k, s_k + m)
// of  a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i  (m >= 2, or m == 0 on idle
// lanes; the first row of the first segment has a = 0 and the last row of
// the last segment c = 0).  Each lane eliminates its segment down and up
// keeping its first unknown as a parameter, the first unknown is folded
// out with the next lane's first row, the one-equation-per-lane system in
// the segment-end unknowns is solved by parallel cyclic reduction over the
// group, and each lane back-substitutes.  On return x holds the solution.
// FULL: every segment but the line's last holds exactly MMAX rows.  Each lane
// then eliminates all MMAX rows with no per-row guard: the rows past a lane's
// m are identity rows (a = c = 0, b = 1, d = 0), decoupled from the line
// because the last segment's last row has c = 0 (its boundary coupling is on
// the right-hand side) and the lanes past the last segment hold nothing else.
template <int L, int MMAX, bool FULL, typename R>
__device__ __forceinline__ void group_solve(R (&a)[MMAX], const R (&b)[MMAX], R (&c)[MMAX],
                                            R (&d)[MMAX], int m_lane, int kseg) {
  const int m = FULL ? MMAX : m_lane;
  R A1 = 0, B1 = 1, C1 = 0, D1 = 0;
  R A2 = 0, C2 = 0, D2 = 0;
  if (m >= 2) {
    {
      const R r = rcp_nr(b[1]);
      a[1] = a[1] * r;
      c[1] = c[1] * r;
      d[1] = d[1] * r;
      A2 = a[1];
      C2 = c[1];
      D2 = d[1];
    }
#pragma unroll
    for (int i = 2; i < MMAX; i++) {
      if (i < m) {
        const R r = rcp_nr(fmar(-a[i], c[i - 1], b[i]));
        d[i] = fmar(-a[i], d[i - 1], d[i]) * r;
        a[i] = -a[i] * a[i - 1] * r;
        c[i] = c[i] * r;
        A2 = a[i];
        C2 = c[i];
        D2 = d[i];
      }
    }
#pragma unroll
    for (int i = MMAX - 3; i >= 1; i--) {
      if (i <= m - 3) {
        d[i] = fmar(-c[i], d[i + 1], d[i]);
        a[i] = fmar(-c[i], a[i + 1], a[i]);
        c[i] = -c[i] * c[i + 1];
      }
    }
    if (m == 2) {
      A1 = a[0];
      B1 = b[0];
      C1 = c[0];
      D1 = d[0];
    } else {
      A1 = a[0];
      B1 = fmar(-c[0], a[1], b[0]);
      C1 = -c[0] * c[1];
      D1 = fmar(-c[0], d[1], d[0]);
    }
  }
  const R A1n = grp_down<L>(A1, 1, kseg, R(0));
  const R B1n = grp_down<L>(B1, 1, kseg, R(1));
  const R C1n = grp_down<L>(C1, 1, kseg, R(0));
  const R D1n = grp_down<L>(D1, 1, kseg, R(0));
  R ea = 0, eb = 1, ec = 0, ed = 0;
  const R i1 = rcp_nr(B1);
  if (m >= 2) {
    const R in = rcp_nr(B1n);
    ea = -A2 * A1 * i1;
    eb = R(1) - A2 * C1 * i1 - C2 * A1n * in;
    ec = -C2 * C1n * in;
    ed = D2 - A2 * D1 * i1 - C2 * D1n * in;
  }
#pragma unroll
  for (int s = 1; s < L; s <<= 1) {
    const R am = grp_up<L>(ea, s, kseg, R(0)), bm = grp_up<L>(eb, s, kseg, R(1));
    const R cm = grp_up<L>(ec, s, kseg, R(0)), dm = grp_up<L>(ed, s, kseg, R(0));
    const R ap = grp_down<L>(ea, s, kseg, R(0)), bp = grp_down<L>(eb, s, kseg, R(1));
    const R cp = grp_down<L>(ec, s, kseg, R(0)), dp = grp_down<L>(ed, s, kseg, R(0));
    const R k1 = ea * rcp_nr(bm), k2 = ec * rcp_nr(bp);
    const R na = -am * k1;
    const R nc = -cp * k2;
    const R nb = fmar(-ap, k2, fmar(-cm, k1, eb));
    const R nd = fmar(-dp, k2, fmar(-dm, k1, ed));
    ea = na;
    eb = nb;
    ec = nc;
    ed = nd;
  }
  const R xe = ed * rcp_nr(eb);
  const R xprev = grp_up<L>(xe, 1, kseg, R(0));
  // (written as selects of computed values: conditional stores of one value
  // to several elements get merged into one indexed store, which keeps the
  // array in scratch memory)
  const R xs = fmar(-C1, xe, fmar(-A1, xprev, D1)) * i1;
  const bool act = m >= 2;
#pragma unroll
  for (int i = 0; i < MMAX; i++) {
    const R mid = fmar(-c[i], xe, fmar(-a[i], xs, d[i]));
    R v = i == 0 ? xs : (i == m - 1 ? xe : mid);
    d[i] = act && i < m ? v : d[i];
  }
}

// The sequential Thomas algorithm in the reference's TDMA operation order
// (douglas_adi.cc:1784-1807) over the whole line, rows 0 and n-1 included,
// for the groups whose line is flagged (seq_line): the partitioned solve is
// stable only for diagonally dominant lines.  The recurrence runs in every
// lane of the group on rows gathered from their owner lanes; the owner keeps
// gamma in a and the forward value, then the solution, in d.  Solved rows
// become identity rows (a = c = 0, b = 1, d = x), which the partitioned
// solve that follows returns unchanged.
template <int L, int MMAX, typename R>
__device__ __forceinline__ void group_thomas(R (&a)[MMAX], R (&b)[MMAX], R (&c)[MMAX],
                                             R (&d)[MMAX], const LinePart &q, int lane, int kseg,
                                             bool seq_line, double x0, double xn) {
  // in FP64 whatever R is: these lines are the ill-conditioned ones
  const int gb = lane - kseg;
  double bet = 1.0, uprev = x0, cprev = 0.0;
  for (int kk = 0; kk < q.p; kk++) {
    const int mk = kk == q.p - 1 ? q.mlast : q.m;
    const int src = gb + kk;
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      if (r < mk) {
        const double ai = __shfl((double)a[r], src, 64), bi = __shfl((double)b[r], src, 64);
        const double ci = __shfl((double)c[r], src, 64), di = __shfl((double)d[r], src, 64);
        const double gam = cprev / bet;
        bet = bi - ai * gam;
        const double u = (di - ai * uprev) / bet;
        if (seq_line && kseg == kk) {
          a[r] = (R)gam;
          d[r] = (R)u;
        }
        cprev = ci;
        uprev = u;
      }
    }
  }
  double gnext = cprev / bet, unext = xn;  // row n-1: gamma, and its value
  for (int kk = q.p - 1; kk >= 0; kk--) {
    const int mk = kk == q.p - 1 ? q.mlast : q.m;
    const int src = gb + kk;
#pragma unroll
    for (int r = MMAX - 1; r >= 0; r--) {
      if (r < mk) {
        const double ui = __shfl((double)d[r], src, 64), gi = __shfl((double)a[r], src, 64);
        const double u = ui - gnext * unext;
        if (seq_line && kseg == kk) d[r] = (R)u;
        gnext = gi;
        unext = u;
      }
    }
  }
  if (seq_line) {
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      a[r] = 0;
      b[r] = 1;
      c[r] = 0;
    }
  }
}

// ----------------------------------------------------------- line kernel
// Tiles that share one 128-byte line: y/z tiles are T consecutive x, so a
// tile narrower than 128 bytes is one of 128 / (T * sizeof(R)) per line.
template <int DIR, int T, typename R>
__device__ __host__ constexpr unsigned tile_group() {
  return (unsigned)line_group(DIR, T, (int)sizeof(R));
}

// Streaming (nontemporal) 16- or 8-byte pair accesses of the field: each
// element of a stage's input and output is touched once, so the hint keeps it
// from displacing the tables in the caches.
typedef double dvec2_t __attribute__((ext_vector_type(2)));
typedef float fvec2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 nt_load2(const double2 *a) {
  const dvec2_t t = __builtin_nontemporal_load(reinterpret_cast<const dvec2_t *>(a));
  return make_double2(t.x, t.y);
}
__device__ __forceinline__ float2 nt_load2(const float2 *a) {
  const fvec2_t t = __builtin_nontemporal_load(reinterpret_cast<const fvec2_t *>(a));
  return make_float2(t.x, t.y);
}
__device__ __forceinline__ void nt_store2(double2 *a, double2 v) {
  dvec2_t t;
  t.x = v.x;
  t.y = v.y;
  __builtin_nontemporal_store(t, reinterpret_cast<dvec2_t *>(a));
}
__device__ __forceinline__ void nt_store2(float2 *a, float2 v) {
  fvec2_t t;
  t.x = v.x;
  t.y = v.y;
  __builtin_nontemporal_store(t, reinterpret_cast<fvec2_t *>(a));
}

// Pair of neighbouring field elements moved as one access.
template <typename R>
struct Pair2;
template <>
struct Pair2<double> {
  using type = double2;
};
template <>
struct Pair2<float> {
  using type = float2;
};

// LDS layout of a line kernel: the field tile(s), the per-line weights, the z
// axis table (z stages with a source term), then the regular lines' solve
// tables (both sides, solve stages), 16-byte aligned.
template <int DIR, int L, int MMAX, int FLAGS, typename R>
struct LdsLayout {
  static constexpr bool SOLVE = (FLAGS & kFlagSolve) != 0;
  static constexpr bool TWO = ((FLAGS & 3) == kRhsDouglas) || !SOLVE;
  static constexpr bool ZT = DIR == 2 && ((FLAGS & kFlagEpiF) || (FLAGS & 3) == kRhsTSZ);
  static constexpr int T = tile_lines(L);
  MTS_RS_HD static size_t wt_off(int lds_len) {
    return (size_t)T * lds_len * sizeof(R) * (TWO ? 2 : 1);
  }
  MTS_RS_HD static size_t zt_off(int lds_len) { return wt_off(lds_len) + (size_t)T * kWtLen * sizeof(double); }
  MTS_RS_HD static size_t reg_off(int lds_len, int n2) {
    const size_t e = zt_off(lds_len) + (ZT ? 2 * (size_t)n2 * sizeof(double) : 0);
    return (e + 15) / 16 * 16;
  }
  MTS_RS_HD static size_t bytes(int lds_len, int n2, bool reg) {
    return reg_off(lds_len, n2) + (reg && SOLVE ? 2 * (size_t)reg_side_len(L, MMAX) * sizeof(R) : 0);
  }
  // the regular-tile kernel: tile(s), weights, then the solve tables of the
  // sides present (reg_side elements each, two sides), no z table
  MTS_RS_HD static size_t reg_bytes(int lds_len, int reg_side, int reg_sides) {
    const size_t e = (zt_off(lds_len) + 15) / 16 * 16;
    return e + ((size_t)reg_sides * reg_side * sizeof(R) + 15) / 16 * 16;
  }
};

// R: the field type, double (FP64 form) or float (FP32 form: fields, LDS
// tiles and the line solve in FP32, rows formed in FP64).
// Tile (bx, by) of workgroup blockIdx.x (tile lists, see tile_lists) or of
// (blockIdx.x, blockIdx.y) (the whole grid).  y/z tiles narrower than a
// 128-byte line (T * sizeof(R) < 128): the G tiles of each line go to one XCD
// (workgroups go to the 8 XCDs in turn), dispatched one after the other, so
// the line is read from HBM once and the other parts come from that XCD's
// L2.  (Otherwise neighbouring tiles stay on different XCDs at the same time:
// giving each XCD a contiguous range of tiles instead is slower in every
// direction.)  Tile lists come in that order already (xcd_group_tiles).
template <int DIR, int T, typename R>
__device__ __forceinline__ void tile_coords(const int *list, int &bx, int &by) {
  constexpr unsigned G = tile_group<DIR, T, R>();
  if (list) {
    const int e = list[blockIdx.x];
    bx = tile_bx(e);
    by = tile_k(e) - 1;
    return;
  }
  const unsigned nb = gridDim.x * gridDim.y;
  const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
  unsigned v = b;
  if (G > 1 && b < nb / (8 * G) * (8 * G)) v = ((b >> 3) / G * 8 + (b & 7)) * G + (b >> 3) % G;
  bx = (int)(v % gridDim.x);
  by = (int)(v / gridDim.x);
}

// The tile walk of thread tid, formed from inputs the compiler must treat as
// new: the per-pair addresses and LDS slots of one phase (load, LDS store,
// write-back) are then recomputed there (a few integer operations) instead
// of being kept in registers through the solve.
template <int DIR, int T>
__device__ __forceinline__ TileWalk<DIR, T> fresh_walk(const LineArgs &p, int j0, int k, int tid) {
  asm volatile("" : "+v"(j0), "+v"(k), "+v"(tid));
  return TileWalk<DIR, T>(p, j0, k, tid);
}

// Douglas x sweep (kFlagPre1): stage the tile of u^n as the stage-1
// right-hand side.  CHUNK pairs at a time, each pair loading its four y / z
// neighbour pairs, its classification and its x tables, all in flight
// together: in the general kernel 3 with 16-row segments (its solve needs
// more registers than the chunk anyway), else 1 (3 waves per SIMD); in the
// regular kernel 1, which keeps it within 128 VGPRs at 4 workgroups per CU
// (2 at 3 workgroups per CU measured 130 us against 107 at 257^3).  End rows come from tile_end_val; pairs of lines past
// the grid or on its boundary are not staged (those lines are not solved).
template <int T, int CHUNK>
__device__ __forceinline__ void pre1_stage_tile(const LineArgs &p, int j0, int k, int tid,
                                                double *tA) {
  const TileWalk<0, T> walk(p, j0, k, tid);
  const Pre1Line ln = pre1_line(p.g, walk.j < p.g.n[1] ? walk.j : 0, k);
#pragma unroll 1
  for (int u0 = 0; u0 < kTilePairsPerThread; u0 += CHUNK) {
    Pre1Pair e[CHUNK];
    WalkPair tp[CHUNK];
#pragma unroll
    for (int c = 0; c < CHUNK; c++) {
      tp[c] = walk.at(p, u0 + c);
      if (u0 + c >= kTilePairsPerThread) tp[c].ld0 = tp[c].ld1 = false;  // (last chunk)
      if ((tp[c].ld0 && tp[c].w0) || (tp[c].ld1 && tp[c].w1)) pre1_load(p, tp[c].gi, tp[c].r, e[c]);
    }
#pragma unroll
    for (int c = 0; c < CHUNK; c++) {
      if (tp[c].ld0 && tp[c].w0) tA[tp[c].sl0] = pre1_value(p, ln, e[c], tp[c].gi, 0);
      if (tp[c].ld1 && tp[c].w1) tA[tp[c].sl1] = pre1_value(p, ln, e[c], tp[c].gi, 1);
    }
  }
}

// One tile of one line stage: every row built from the tables (regular rows
// and MIB blocks), the general partitioned solve.
template <int DIR, int L, int MMAX, int FLAGS, typename R>
__device__ __forceinline__ void line_body(const LineArgs &p, int bx, int by) {
  using V2 = typename Pair2<R>::type;
  constexpr int RHS = FLAGS & 3;
  constexpr bool SOLVE = (FLAGS & kFlagSolve) != 0;
  constexpr bool HAS_S = (RHS == kRhsDouglas);
  constexpr bool TWO = HAS_S || !SOLVE;  // second LDS tile: S, or the output of explicit stages
  // streaming field accesses in the FP64 Douglas stages (as line_kernel_reg)
  constexpr bool NT = (HAS_S || (DIR == 0 && (FLAGS & kFlagPre1) != 0)) && sizeof(R) == sizeof(double);
  constexpr bool SEQ = SOLVE && (FLAGS & kFlagSeq) != 0;
  constexpr bool FULL = (FLAGS & kFlagFull) != 0;
  constexpr int LPW = 64 / L;          // lines per wavefront
  constexpr int T = tile_lines(L);     // lines per tile (all solved at once)
  const DevGeom &g = p.g;
  const LinePart &q = p.q;
  const int n = g.n[DIR];
  const int nj = LineMap<DIR>::nj(g);
  const int j0 = bx * T;
  const int k = by + 1;
  if (j0 >= nj) return;  // the whole-grid launch rounds its width up to a multiple of G

  extern __shared__ double smem[];
  R *tA = reinterpret_cast<R *>(smem);
  R *tB = tA + T * q.lds_len;
  const R *fA = static_cast<const R *>(p.A);
  const R *fS = static_cast<const R *>(p.S);
  R *fO = static_cast<R *>(p.out);

  // stage the tile: loads of element pairs, kChunk of them in flight per
  // thread (all of a thread's pairs), then their LDS stores
  constexpr int kChunk = kTileChunk;
  const int tid = threadIdx.x;
  // The small tables that join the tile (LDS after the field tiles, whose
  // size is a multiple of 8 bytes: T is even): the per-line operator weights
  // (LineWeights); for z lines with a source term, sin z and cos z of the
  // rows; for in_analytic stages, the lines' end rows (2 per line), which the
  // tile walk does not load.  Their global loads go out before the tile's,
  // so the two latencies overlap.
  constexpr bool ZT = DIR == 2 && ((FLAGS & kFlagEpiF) || RHS == kRhsTSZ);
  constexpr int kZPre = 3;  // z-table entries a thread loads ahead of the tile
  const double wt_v = tid < T * kWtLen ? line_weight_value<DIR>(p, j0, k, tid) : 0.0;
  double zt_v[kZPre];
#pragma unroll
  for (int i = 0; i < kZPre; i++) {
    const int e = tid + i * kBlock;
    zt_v[i] = ZT && e < 2 * n ? ztab_value(g, e) : 0.0;
  }
  // this lane's line: its table entries and boundary values, also loaded
  // ahead of the tile
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int kseg = lane & (L - 1);
  const int l = wave * LPW + lane / L;
  const int j = j0 + l;
  const bool valid = j >= 1 && j <= nj - 2;
  const LineMeta lm = line_meta(p.mib, k * nj + j, valid);
  const LaneSeg sg = lane_seg(q, kseg, valid);
  double x0 = 0.0, xn = 0.0;
  if (valid) line_ends<DIR>(p, j, k, x0, xn);
  const bool end_row = p.in_analytic && tid < 2 * T;
  const int end_sl = end_row ? tile_end_slot<DIR>(p, j0, tid) : -1;
  const double end_v = end_sl >= 0 ? tile_end_val<DIR>(p, j0, k, tid) : 0.0;

  constexpr bool PRE1 = DIR == 0 && (FLAGS & kFlagPre1) != 0;
  if constexpr (PRE1) {
    pre1_stage_tile<T, (MMAX >= 16 ? 3 : 1)>(p, j0, k, tid, reinterpret_cast<double *>(tA));
  } else {
  const TileWalk<DIR, T> walk(p, j0, k, tid);
#pragma unroll 1
  for (int u0 = 0; u0 < kTilePairsPerThread; u0 += kChunk) {
    V2 va[kChunk], vb[kChunk];
    WalkPair tp[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; c++) {
      tp[c] = walk.at(p, u0 + c);
      va[c].x = va[c].y = R(0);
      vb[c].x = vb[c].y = R(0);
      if (tp[c].ld0 || tp[c].ld1) {  // (ld1 implies the first element exists)
        if constexpr (NT) {
          va[c] = nt_load2(reinterpret_cast<const V2 *>(fA + tp[c].gi));
          if (HAS_S) vb[c] = nt_load2(reinterpret_cast<const V2 *>(fS + tp[c].gi));
        } else {
          va[c] = *reinterpret_cast<const V2 *>(fA + tp[c].gi);
          if (HAS_S) vb[c] = *reinterpret_cast<const V2 *>(fS + tp[c].gi);
        }
      }
    }
    const TileWalk<DIR, T> walk_st = fresh_walk<DIR, T>(p, j0, k, tid);
#pragma unroll
    for (int c = 0; c < kChunk; c++) {
      const WalkPair t = walk_st.at(p, u0 + c);
      if (t.ld0) {
        tA[t.sl0] = va[c].x;
        if (HAS_S) tB[t.sl0] = vb[c].x;
      }
      if (t.ld1) {
        tA[t.sl1] = va[c].y;
        if (HAS_S) tB[t.sl1] = vb[c].y;
      }
    }
  }
  }
  double *wt = reinterpret_cast<double *>(tA + (size_t)T * q.lds_len * (TWO ? 2 : 1));
  double *zt = ZT ? wt + T * kWtLen : nullptr;
  if (tid < T * kWtLen) wt[tid] = wt_v;
  if (ZT) {
#pragma unroll
    for (int i = 0; i < kZPre; i++) {
      const int e = tid + i * kBlock;
      if (e < 2 * n) zt[e] = zt_v[i];
    }
#pragma unroll 1
    for (int e = tid + kZPre * kBlock; e < 2 * n; e += kBlock) ztab_entry(g, e, zt);
  }
  if (end_sl >= 0) tA[end_sl] = (R)end_v;
  __syncthreads();

  {
    R a[MMAX], b[MMAX], c[MMAX], d[MMAX];
    line_rows<DIR, MMAX, FLAGS>(p, j, k, lm, tA + l * q.lds_len, tB + l * q.lds_len,
                                wt + l * kWtLen, zt, sg, a, b, c, d);
    if (SOLVE) {
      if (SEQ) {
        if (__any(lm.seq)) group_thomas<L, MMAX, R>(a, b, c, d, q, lane, kseg, lm.seq, x0, xn);
      }
      // boundary rows are known: move their couplings to the right-hand side
      const bool first = sg.m > 0 && sg.k == 0;
      d[0] = first ? (R)((double)d[0] - (double)a[0] * x0) : d[0];
      a[0] = first ? R(0) : a[0];
      const bool lastseg = sg.m > 0 && sg.k == q.p - 1;
#pragma unroll
      for (int r = 0; r < MMAX; r++) {
        const bool hit = lastseg && r == sg.m - 1;
        d[r] = hit ? (R)((double)d[r] - (double)c[r] * xn) : d[r];
        c[r] = hit ? R(0) : c[r];
      }
      group_solve<L, MMAX, FULL, R>(a, b, c, d, sg.m, kseg);
    }
    if (valid)
      line_write<DIR, MMAX, FLAGS>(p, j, k, lm, sg, d, x0, xn,
                                   (TWO && !HAS_S ? tB : tA) + l * q.lds_len, zt);
  }
  __syncthreads();

  const R *tO = TWO && !HAS_S ? tB : tA;
  const TileWalk<DIR, T> walk_wb = fresh_walk<DIR, T>(p, j0, k, tid);
#pragma unroll 1
  for (int u0 = 0; u0 < kTilePairsPerThread; u0 += kChunk) {
    V2 vo[kChunk];
    WalkPair tp[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; c++) {
      tp[c] = walk_wb.at(p, u0 + c);
      vo[c].x = tp[c].w0 ? tO[tp[c].sl0] : R(0);
      vo[c].y = tp[c].w1 ? tO[tp[c].sl1] : R(0);
    }
#pragma unroll
    for (int c = 0; c < kChunk; c++) {
      const WalkPair &t = tp[c];
      if (t.w0 && t.w1) {
        if constexpr (NT) nt_store2(reinterpret_cast<V2 *>(fO + t.gi), vo[c]);
        else *reinterpret_cast<V2 *>(fO + t.gi) = vo[c];
      } else {
        if (t.w0) fO[t.gi] = vo[c].x;
        if (t.w1) fO[t.gi + 1] = vo[c].y;
      }
    }
  }
}

template <int DIR, int L, int MMAX, int FLAGS, typename R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void line_kernel(
    LineArgs p) {
  int bx, by;
  tile_coords<DIR, tile_lines(L), R>(p.split ? p.tiles[1] : nullptr, bx, by);
  line_body<DIR, L, MMAX, FLAGS, R>(p, bx, by);
}
// The slot of workgroup b in a launch over several instances whose first
// workgroups are first[0..count) (ascending, first[0] = 0; a slot without
// workgroups shares its start with the next one), and b's tile in it.
__device__ __forceinline__ int2 multi_slot(const int *first, int count, unsigned b) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if ((unsigned)first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return make_int2(lo, (int)(b - (unsigned)first[lo]));
}
// The general tiles of several instances' line stage in one launch:
// workgroup b solves tile y of the list of the instance in slot x
// (multi_slot), whose stage arguments are args[x] (in device memory).
template <int DIR, int L, int MMAX, int FLAGS, typename R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2))) void line_kernel_multi(
    const LineArgs *args, const int *first, int count) {
  const int2 w = multi_slot(first, count, blockIdx.x);
  const LineArgs &p = args[w.x];
  const int e = p.tiles[1][w.y];
  line_body<DIR, L, MMAX, FLAGS, R>(p, tile_bx(e), tile_k(e) - 1);
}

// What one lane needs of one tile besides the field: the tile's place, the
// lane's line (interior or not, side), its boundary values and the line's
// source factors, the thread's entry of the tile's weight table and, for
// in_analytic stages, of its end rows.  Gathered for the next tile while the
// current one is solved.
template <int DIR>
struct RegTileMeta {
  int j0, k;
  bool valid, out0;
  double x0, xn, wt_v, end_v;
  int end_sl;
  LineSource<DIR> src;
};

template <int DIR, int L>
__device__ __forceinline__ RegTileMeta<DIR> reg_tile_meta(const LineArgs &p, int e,
                                                          const double *zt) {
  constexpr int T = tile_lines(L);
  constexpr int LPW = 64 / L;
  const int tid = threadIdx.x;
  const int l = (tid >> 6) * LPW + (tid & 63) / L;
  const int nj = LineMap<DIR>::nj(p.g);
  RegTileMeta<DIR> m{};
  m.j0 = tile_bx(e) * T;
  m.k = tile_k(e);
  const int j = m.j0 + l;
  m.valid = j >= 1 && j <= nj - 2;
  m.out0 = m.valid && p.mib.cls0[m.k * nj + j] != 0;
  m.x0 = m.xn = 0.0;
  if (m.valid) line_ends<DIR>(p, j, m.k, m.x0, m.xn);
  m.wt_v = tid < T * kWtLen ? line_weight_value<DIR>(p, m.j0, m.k, tid) : 0.0;
  m.end_sl = p.in_analytic && tid < 2 * T ? tile_end_slot<DIR>(p, m.j0, tid) : -1;
  m.end_v = m.end_sl >= 0 ? tile_end_val<DIR>(p, m.j0, m.k, tid) : 0.0;
  if (m.valid) m.src = LineSource<DIR>(p.P, p.g, j, m.k, zt);
  return m;
}

// Solve stage, regular tiles (every line regular: no MIB block, a constant
// coefficient on its side): the right-hand-side-only solve with the
// precomputed tables, one tile per workgroup.  None of the general row, block
// and solve code is here, so the kernel needs far fewer registers and LDS
// than line_kernel and four of its workgroups share a CU: while some load
// their tile, others solve or store theirs.
template <int DIR, int L, int MMAX, int FLAGS, typename R>
__device__ __forceinline__ void reg_body(const LineArgs &p, int tile) {
  using V2 = typename Pair2<R>::type;
  constexpr int RHS = FLAGS & 3;
  constexpr bool HAS_S = (RHS == kRhsDouglas);
  constexpr bool FULL = (FLAGS & kFlagFull) != 0;
  constexpr int T = tile_lines(L);
  constexpr int kChunk = kTileChunk;
  using LL = LdsLayout<DIR, L, MMAX, FLAGS, R>;
  const DevGeom &g = p.g;
  const LinePart &q = p.q;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int kseg = lane & (L - 1);
  const int l = (tid >> 6) * (64 / L) + lane / L;

  extern __shared__ double smem[];
  R *tA = reinterpret_cast<R *>(smem);
  R *tB = tA + T * q.lds_len;
  double *wt = reinterpret_cast<double *>(reinterpret_cast<char *>(smem) + LL::wt_off(q.lds_len));
  R *lreg = reinterpret_cast<R *>(reinterpret_cast<char *>(smem) + (LL::zt_off(q.lds_len) + 15) / 16 * 16);
  const R *fA = static_cast<const R *>(p.A);
  const R *fS = static_cast<const R *>(p.S);
  R *fO = static_cast<R *>(p.out);

  // this tile: the lane's line, boundary values and source factors (the z
  // stages read sin z / cos z from the axis tables, not an LDS copy), the
  // solve table of the sides present, then the field, all loads in flight
  // together
  const RegTileMeta<DIR> cur = reg_tile_meta<DIR, L>(p, tile, nullptr);
  {
    // the solve tables straight into LDS (16-byte LDS-DMA loads: lane i of
    // a wavefront writes 16 bytes at the instruction's LDS base + 16 i); the
    // host pads them to whole 16-byte units
    const char *greg = static_cast<const char *>(p.reg[FULL ? 1 : 0]);
    const int units = (p.reg_sides * p.reg_side * (int)sizeof(R) + 15) / 16;
    const int wave = tid >> 6;
    for (int u0 = wave * 64; u0 < units; u0 += kBlock) {
      if (u0 + lane < units)
        __builtin_amdgcn_global_load_lds(
            greg + (size_t)(u0 + lane) * 16,
            (__attribute__((address_space(3))) void *)(reinterpret_cast<char *>(lreg) +
                                                         (size_t)u0 * 16),
            16, 0, 0);
    }
  }
  constexpr bool PRE1 = DIR == 0 && (FLAGS & kFlagPre1) != 0;
  // FP64 Douglas stages stream their field accesses past the caches
  // (measured at 257^3: 3-4% faster y, z and fused x stages; the LOD stages
  // lose as much, and the FP32 ones, whose y/z tiles share 128-byte lines
  // through the L2, 13%, so they keep plain accesses)
  constexpr bool NT = (HAS_S || PRE1) && sizeof(R) == sizeof(double);
  if constexpr (PRE1) {
    pre1_stage_tile<T, 1>(p, cur.j0, cur.k, tid, reinterpret_cast<double *>(tA));
  } else {
  const TileWalk<DIR, T> walk(p, cur.j0, cur.k, tid);
  V2 va[kChunk], vb[kChunk];
#pragma unroll
  for (int c = 0; c < kChunk; c++) {
    const WalkPair w = walk.at(p, c);
    va[c].x = va[c].y = R(0);
    vb[c].x = vb[c].y = R(0);
    if (w.ld0 || w.ld1) {
      if constexpr (NT) {
        va[c] = nt_load2(reinterpret_cast<const V2 *>(fA + w.gi));
        if (HAS_S) vb[c] = nt_load2(reinterpret_cast<const V2 *>(fS + w.gi));
      } else {
        va[c] = *reinterpret_cast<const V2 *>(fA + w.gi);
        if (HAS_S) vb[c] = *reinterpret_cast<const V2 *>(fS + w.gi);
      }
    }
  }
  const TileWalk<DIR, T> walk_st = fresh_walk<DIR, T>(p, cur.j0, cur.k, tid);
#pragma unroll
  for (int c = 0; c < kChunk; c++) {
    const WalkPair w = walk_st.at(p, c);
    if (w.ld0) {
      tA[w.sl0] = va[c].x;
      if (HAS_S) tB[w.sl0] = vb[c].x;
    }
    if (w.ld1) {
      tA[w.sl1] = va[c].y;
      if (HAS_S) tB[w.sl1] = vb[c].y;
    }
  }
  }
  if (tid < T * kWtLen) wt[tid] = cur.wt_v;
  if (cur.end_sl >= 0) tA[cur.end_sl] = (R)cur.end_v;
  __builtin_amdgcn_s_waitcnt(0);  // the LDS-DMA loads are in LDS
  __syncthreads();

  {
    const int j = cur.j0 + l;
    const LaneSeg sg = lane_seg(q, kseg, cur.valid);
    R d[MMAX];
    regular_rhs<DIR, MMAX, FLAGS>(p, j, cur.k, cur.out0, tA + l * q.lds_len, tB + l * q.lds_len,
                                  wt + l * kWtLen, cur.src, sg, d);
    const double *w = wt + l * kWtLen + (cur.out0 ? 0 : 3);
    const R rt = (R)p.theta_dt;
    const R a_first = -rt * (R)w[0], c_last = -rt * (R)w[2];
    const bool first = sg.m > 0 && sg.k == 0;
    d[0] = first ? (R)((double)d[0] - (double)a_first * cur.x0) : d[0];
    const bool lastseg = sg.m > 0 && sg.k == q.p - 1;
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      const bool hit = lastseg && r == sg.m - 1;
      d[r] = hit ? (R)((double)d[r] - (double)c_last * cur.xn) : d[r];
    }
    const R *tb = lreg + (cur.out0 ? 0 : p.reg_side);
    const R *rows = tb + reg_rows_off(kseg == q.p - 1 ? 1 : 0, MMAX);
    const R *sc = tb + reg_scal_off(kseg, MMAX);
    R D1, D2;
    reg_forward<MMAX, FULL, R>(rows, sc, d, sg.m, D1, D2);
    const R D1n = grp_down<L>(D1, 1, kseg, R(0));
    R ed = reg_fold<R>(sc, FULL ? MMAX : sg.m, D1, D2, D1n);
#pragma unroll
    for (int lv = 0, s = 1; s < L; lv++, s <<= 1) {
      const R dm = grp_up<L>(ed, s, kseg, R(0)), dp = grp_down<L>(ed, s, kseg, R(0));
      ed = reg_level<R>(sc, lv, ed, dm, dp);
    }
    const R xe = ed * sc[kRegQ];
    const R xprev = grp_up<L>(xe, 1, kseg, R(0));
    reg_back<MMAX, FULL, R>(rows, sc, d, sg.m, D1, xe, xprev);
    if (cur.valid) {
      LineMeta lm;
      lm.bb = lm.be = 0;
      lm.out0 = cur.out0;
      lm.seq = false;
      line_write<DIR, MMAX, FLAGS>(p, j, cur.k, lm, sg, d, cur.x0, cur.xn, tA + l * q.lds_len,
                                   cur.src);
    }
  }
  __syncthreads();

  const TileWalk<DIR, T> walk_wb = fresh_walk<DIR, T>(p, cur.j0, cur.k, tid);
#pragma unroll
  for (int c = 0; c < kChunk; c++) {
    const WalkPair w = walk_wb.at(p, c);
    V2 vo;
    vo.x = w.w0 ? tA[w.sl0] : R(0);
    vo.y = w.w1 ? tA[w.sl1] : R(0);
    if (w.w0 && w.w1) {
      if constexpr (NT) nt_store2(reinterpret_cast<V2 *>(fO + w.gi), vo);
      else *reinterpret_cast<V2 *>(fO + w.gi) = vo;
    } else {
      if (w.w0) fO[w.gi] = vo.x;
      if (w.w1) fO[w.gi + 1] = vo.y;
    }
  }
}
template <int DIR, int L, int MMAX, int FLAGS, typename R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void line_kernel_reg(LineArgs p) {
  reg_body<DIR, L, MMAX, FLAGS, R>(p, p.tiles[0][blockIdx.x]);
}
// The regular tiles of several instances' line stage in one launch (as
// line_kernel_multi).
template <int DIR, int L, int MMAX, int FLAGS, typename R>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void line_kernel_reg_multi(
    const LineArgs *args, const int *first, int count) {
  const int2 w = multi_slot(first, count, blockIdx.x);
  const LineArgs &p = args[w.x];
  reg_body<DIR, L, MMAX, FLAGS, R>(p, p.tiles[0][w.y]);
}
// A solve stage with tile lists runs the regular tiles with line_kernel_reg
// and the others with line_kernel; any other stage runs line_kernel over the
// whole grid.
template <int DIR, int L, int MMAX, int FLAGS, typename R>
hipError_t launch_line_t(const LineArgs &a, hipStream_t st) {
  const DevGeom &g = a.g;
  const int nj = DIR == 0 ? g.n[1] : g.n[0];
  const int nk = DIR == 2 ? g.n[1] : g.n[2];
  constexpr int T = tile_lines(L);
  constexpr bool SOLVE = (FLAGS & kFlagSolve) != 0;
  constexpr bool FULL = (FLAGS & kFlagFull) != 0;
  using LL = LdsLayout<DIR, L, MMAX, FLAGS, R>;
  if (SOLVE && a.split) {
    if (a.ntiles[1] > 0) {
      hipLaunchKernelGGL((line_kernel<DIR, L, MMAX, FLAGS, R>), dim3(a.ntiles[1]), dim3(kBlock),
                         LL::bytes(a.q.lds_len, g.n[2], false), st, a);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    if (a.ntiles[0] > 0) {
      if (!a.reg[FULL ? 1 : 0]) return hipErrorInvalidValue;
      hipLaunchKernelGGL((line_kernel_reg<DIR, L, MMAX, FLAGS, R>), dim3(a.ntiles[0]), dim3(kBlock),
                         LL::reg_bytes(a.q.lds_len, a.reg_side, a.reg_sides), st, a);
      return hipGetLastError();
    }
    return hipSuccess;
  }
  LineArgs b = a;
  b.split = 0;
  const size_t shm = LL::bytes(a.q.lds_len, g.n[2], false);
  int gx = (nj + T - 1) / T;
  constexpr int G = (int)tile_group<DIR, T, R>();
  gx = (gx + G - 1) / G * G;  // tiles in groups of G (see tile_coords)
  dim3 grid(gx, nk - 2);
  hipLaunchKernelGGL((line_kernel<DIR, L, MMAX, FLAGS, R>), grid, dim3(kBlock), shm, st, b);
  return hipGetLastError();
}

// One group of a multi-instance line stage: the general tiles (regular
// false) or the regular tiles of the instances whose stage arguments are
// hargs[0..nargs) (host copies, for the LDS size) and dargs (device), nwork
// workgroups, those of slot i from dfirst[i].
template <int DIR, int L, int MMAX, int FLAGS, typename R>
hipError_t launch_multi_t(bool regular, const LineArgs *hargs, int nargs, const LineArgs *dargs,
                          const int *dfirst, int nwork, hipStream_t st) {
  using LL = LdsLayout<DIR, L, MMAX, FLAGS, R>;
  constexpr bool FULL = (FLAGS & kFlagFull) != 0;
  size_t shm = 0;
  for (int i = 0; i < nargs; i++) {
    const LineArgs &a = hargs[i];
    if (regular && !a.reg[FULL ? 1 : 0]) return hipErrorInvalidValue;
    const size_t b = regular ? LL::reg_bytes(a.q.lds_len, a.reg_side, a.reg_sides)
                             : LL::bytes(a.q.lds_len, a.g.n[2], false);
    shm = b > shm ? b : shm;
  }
  if (nwork <= 0) return hipSuccess;
  if (regular)
    hipLaunchKernelGGL((line_kernel_reg_multi<DIR, L, MMAX, FLAGS, R>), dim3(nwork), dim3(kBlock),
                       shm, st, dargs, dfirst, nargs);
  else
    hipLaunchKernelGGL((line_kernel_multi<DIR, L, MMAX, FLAGS, R>), dim3(nwork), dim3(kBlock), shm,
                       st, dargs, dfirst, nargs);
  return hipGetLastError();
}
template <int FLAGS, typename R>
hipError_t multi_flags(int mmax, bool full, bool regular, const LineArgs *hargs, int nargs,
                       const LineArgs *dargs, const int *dwork, int nwork, hipStream_t st) {
  constexpr int D = MTS_LINE_DIR, LL = MTS_LINE_L;
  constexpr bool CAN_FULL = (FLAGS & kFlagSolve) && !(FLAGS & kFlagSeq);
  constexpr int FF = CAN_FULL ? (FLAGS | kFlagFull) : FLAGS;
  full = CAN_FULL && full;
#define MTS_MULTI_M(M)                                                                         \
  if (mmax == M)                                                                               \
    return full ? launch_multi_t<D, LL, M, FF, R>(regular, hargs, nargs, dargs, dwork, nwork, st) \
                : launch_multi_t<D, LL, M, FLAGS, R>(regular, hargs, nargs, dargs, dwork, nwork, st);
  MTS_MULTI_M(4)
  MTS_MULTI_M(8)
  MTS_MULTI_M(16)
#undef MTS_MULTI_M
  return hipErrorInvalidValue;
}

// Solve stages whose split has every segment but the last at exactly mmax
// rows run the unguarded solve (kFlagFull; the sequential-Thomas variant keeps
// the guarded one).
template <int FLAGS, typename R>
hipError_t launch_flags(int mmax, const LineArgs &a, hipStream_t st) {
  constexpr int D = MTS_LINE_DIR, LL = MTS_LINE_L;
  constexpr bool CAN_FULL = (FLAGS & kFlagSolve) && !(FLAGS & kFlagSeq);
  constexpr int FF = CAN_FULL ? (FLAGS | kFlagFull) : FLAGS;
  const bool full = CAN_FULL && a.q.m == mmax;
  if (mmax == 4)
    return full ? launch_line_t<D, LL, 4, FF, R>(a, st) : launch_line_t<D, LL, 4, FLAGS, R>(a, st);
  if (mmax == 8)
    return full ? launch_line_t<D, LL, 8, FF, R>(a, st) : launch_line_t<D, LL, 8, FLAGS, R>(a, st);
  if (mmax == 16)
    return full ? launch_line_t<D, LL, 16, FF, R>(a, st) : launch_line_t<D, LL, 16, FLAGS, R>(a, st);
  return hipErrorInvalidValue;
}

}  // namespace

// Each object holds one direction, lane-group size and field type (its entry
// points end in _0 for FP64, _1 for FP32).
#define MTS_CAT2(a, b, c, d) a##b##_##c##_##d
#define MTS_CAT(a, b, c, d) MTS_CAT2(a, b, c, d)
#define MTS_ENTRY MTS_CAT(launch_line_dl_, MTS_LINE_DIR, MTS_LINE_L, MTS_LINE_F32)

// The sequential-solve variant is compiled separately so that its register
// needs do not lower the occupancy of the common case (no such line).
#define MTS_LINE_CASE(F, R)                                                  \
  case (F):                                                                  \
    return seq ? launch_flags<(F) | kFlagSeq, R>(a.q.mmax, a, s)            \
               : launch_flags<(F), R>(a.q.mmax, a, s);

#if MTS_LINE_MULTI
#define MTS_MULTI_ENTRY MTS_CAT(launch_multi_dl_, MTS_LINE_DIR, MTS_LINE_L, MTS_LINE_F32)
#define MTS_MULTI_CASE(F, R)                                                                 \
  case (F):                                                                                  \
    return seq ? multi_flags<(F) | kFlagSeq, R>(mmax, full, regular, hargs, nargs, dargs, dwork, \
                                                nwork, s)                                    \
               : multi_flags<(F), R>(mmax, full, regular, hargs, nargs, dargs, dwork, nwork, s);

// The solve stages of the split (general and regular tile lists) only.
hipError_t MTS_MULTI_ENTRY(int flags, bool seq, bool f32, int mmax, bool full, bool regular,
                           const LineArgs *hargs, int nargs, const LineArgs *dargs,
                           const int *dwork, int nwork, hipStream_t s) {
  if (f32 != (MTS_LINE_F32 != 0)) return hipErrorInvalidValue;
#if MTS_LINE_F32
  switch (flags) {
    MTS_MULTI_CASE(kRhsCopy | kFlagSolve, float)
    MTS_MULTI_CASE(kRhsDouglas | kFlagSolve, float)
    MTS_MULTI_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump, float)
    MTS_MULTI_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF, float)
    MTS_MULTI_CASE(kRhsCN | kFlagSolve | kFlagLhsJump, float)
    MTS_MULTI_CASE(kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF, float)
    MTS_MULTI_CASE(kRhsTSZ | kFlagSolve | kFlagLhsJump, float)
    default: return hipErrorInvalidValue;
  }
#else
  switch (flags) {
    MTS_MULTI_CASE(kRhsCopy | kFlagSolve, double)
#if MTS_LINE_DIR == 0
    MTS_MULTI_CASE(kRhsCopy | kFlagSolve | kFlagPre1, double)
#endif
    MTS_MULTI_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump, double)
    MTS_MULTI_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF, double)
    MTS_MULTI_CASE(kRhsDouglas | kFlagSolve, double)
    MTS_MULTI_CASE(kRhsCN | kFlagSolve | kFlagLhsJump, double)
    MTS_MULTI_CASE(kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF, double)
    MTS_MULTI_CASE(kRhsTSZ | kFlagSolve | kFlagLhsJump, double)
    default: return hipErrorInvalidValue;
  }
#endif
}

#else
hipError_t MTS_ENTRY(int flags, bool seq, bool f32, const LineArgs &a, hipStream_t s) {
  if (f32 != (MTS_LINE_F32 != 0)) return hipErrorInvalidValue;
#if MTS_LINE_F32
  // the FP32 form: LOD, TS and Douglas-ADI (its stage 1 outside the x sweep)
  switch (flags) {
    MTS_LINE_CASE(kRhsCopy | kFlagSolve, float)
    MTS_LINE_CASE(kRhsDouglas | kFlagSolve, float)
    MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump, float)
    MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF, float)
    MTS_LINE_CASE(kRhsCN | kFlagSolve | kFlagLhsJump, float)
    MTS_LINE_CASE(kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF, float)
    MTS_LINE_CASE(kRhsTSZ | kFlagSolve | kFlagLhsJump, float)
    case kRhsCN: return launch_flags<kRhsCN, float>(a.q.mmax, a, s);
    default: return hipErrorInvalidValue;
  }
#else
  switch (flags) {
    MTS_LINE_CASE(kRhsCopy | kFlagSolve, double)
#if MTS_LINE_DIR == 0
    MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagPre1, double)
#endif
    MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump, double)
    MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF, double)
    MTS_LINE_CASE(kRhsDouglas | kFlagSolve, double)
    MTS_LINE_CASE(kRhsCN | kFlagSolve | kFlagLhsJump, double)
    MTS_LINE_CASE(kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF, double)
    MTS_LINE_CASE(kRhsTSZ | kFlagSolve | kFlagLhsJump, double)
    case kRhsCN: return launch_flags<kRhsCN, double>(a.q.mmax, a, s);
    default: return hipErrorInvalidValue;
  }
#endif
}

#endif

}  // namespace mts
