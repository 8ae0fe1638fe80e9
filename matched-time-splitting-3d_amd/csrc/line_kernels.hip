// gfx950 line-stage kernels (see kernels.hip for the overview).  This file
// is compiled once per line direction and lane-group size
// (-DMTS_LINE_DIR=0|1|2 -DMTS_LINE_L=16|32|64), so the many template
// instances build in parallel; each object defines one
// launch_line_dl_<dir>_<L> entry used by launch_line.
#include <hip/hip_runtime.h>

#include "device_core.hpp"
#include "kernels.hpp"

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
  return (DIR != 0 && T * sizeof(R) < 128) ? 128u / (T * sizeof(R)) : 1u;
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

// R: the field type, double (FP64 form) or float (FP32 form: fields, LDS
// tiles and the line solve in FP32, rows formed in FP64).
template <int DIR, int L, int MMAX, int FLAGS, typename R>
__global__ __launch_bounds__(kBlock) void line_kernel(LineArgs p) {
  using V2 = typename Pair2<R>::type;
  constexpr int RHS = FLAGS & 3;
  constexpr bool SOLVE = (FLAGS & kFlagSolve) != 0;
  constexpr bool HAS_S = (RHS == kRhsDouglas);
  constexpr bool TWO = HAS_S || !SOLVE;  // second LDS tile: S, or the output of explicit stages
  constexpr bool SEQ = SOLVE && (FLAGS & kFlagSeq) != 0;
  constexpr bool FULL = (FLAGS & kFlagFull) != 0;
  constexpr int LPW = 64 / L;          // lines per wavefront
  constexpr int T = tile_lines(L);     // lines per tile (all solved at once)
  const DevGeom &g = p.g;
  const LinePart &q = p.q;
  const int n = g.n[DIR];
  const int nj = LineMap<DIR>::nj(g);
  int bx = blockIdx.x, by = blockIdx.y;
  constexpr unsigned G = tile_group<DIR, T, R>();
  if (G > 1) {
    // y/z tiles narrower than a 128-byte line (T * sizeof(R) < 128): place
    // the G tiles of each line on one XCD (workgroups go to the 8 XCDs in
    // turn), dispatched one after the other, so the line is read from HBM
    // once and the other parts come from that XCD's L2.  (Otherwise
    // neighbouring tiles stay on different XCDs at the same time: giving
    // each XCD a contiguous range of tiles instead is slower in every
    // direction.)
    const unsigned nb = gridDim.x * gridDim.y;
    const unsigned b = blockIdx.y * gridDim.x + blockIdx.x;
    unsigned v = b;
    if (b < nb / (8 * G) * (8 * G)) v = ((b >> 3) / G * 8 + (b & 7)) * G + (b >> 3) % G;
    bx = v % gridDim.x;
    by = v / gridDim.x;
  }
  const int j0 = bx * T;
  const int k = by + 1;
  if (j0 >= nj) return;  // the launch rounds the grid width up to a multiple of G

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
        va[c] = *reinterpret_cast<const V2 *>(fA + tp[c].gi);
        if (HAS_S) vb[c] = *reinterpret_cast<const V2 *>(fS + tp[c].gi);
      }
    }
#pragma unroll
    for (int c = 0; c < kChunk; c++) {
      const WalkPair &t = tp[c];
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
#pragma unroll 1
  for (int u0 = 0; u0 < kTilePairsPerThread; u0 += kChunk) {
    V2 vo[kChunk];
    WalkPair tp[kChunk];
#pragma unroll
    for (int c = 0; c < kChunk; c++) {
      tp[c] = walk.at(p, u0 + c);
      vo[c].x = tp[c].w0 ? tO[tp[c].sl0] : R(0);
      vo[c].y = tp[c].w1 ? tO[tp[c].sl1] : R(0);
    }
#pragma unroll
    for (int c = 0; c < kChunk; c++) {
      const WalkPair &t = tp[c];
      if (t.w0 && t.w1) {
        *reinterpret_cast<V2 *>(fO + t.gi) = vo[c];
      } else {
        if (t.w0) fO[t.gi] = vo[c].x;
        if (t.w1) fO[t.gi + 1] = vo[c].y;
      }
    }
  }
}

template <int DIR, int L, int MMAX, int FLAGS, typename R>
hipError_t launch_line_t(const LineArgs &a, hipStream_t st) {
  const DevGeom &g = a.g;
  const int nj = DIR == 0 ? g.n[1] : g.n[0];
  const int nk = DIR == 2 ? g.n[1] : g.n[2];
  constexpr bool TWO = ((FLAGS & 3) == kRhsDouglas) || !(FLAGS & kFlagSolve);
  constexpr int T = tile_lines(L);
  constexpr bool ZT = DIR == 2 && ((FLAGS & kFlagEpiF) || (FLAGS & 3) == kRhsTSZ);
  const size_t shm = (size_t)T * a.q.lds_len * sizeof(R) * (TWO ? 2 : 1) +
                    (size_t)T * kWtLen * sizeof(double) + (ZT ? 2 * (size_t)g.n[2] * sizeof(double) : 0);
  int gx = (nj + T - 1) / T;
  constexpr int G = (int)tile_group<DIR, T, R>();
  gx = (gx + G - 1) / G * G;  // tiles in groups of G (see line_kernel)
  dim3 grid(gx, nk - 2);
  hipLaunchKernelGGL((line_kernel<DIR, L, MMAX, FLAGS, R>), grid, dim3(kBlock), shm, st, a);
  return hipGetLastError();
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

#define MTS_CAT2(a, b, c) a##b##_##c
#define MTS_CAT(a, b, c) MTS_CAT2(a, b, c)
#define MTS_ENTRY MTS_CAT(launch_line_dl_, MTS_LINE_DIR, MTS_LINE_L)

// The sequential-solve variant is compiled separately so that its register
// needs do not lower the occupancy of the common case (no such line).
#define MTS_LINE_CASE(F, R)                                                  \
  case (F):                                                                  \
    return seq ? launch_flags<(F) | kFlagSeq, R>(a.q.mmax, a, s)            \
               : launch_flags<(F), R>(a.q.mmax, a, s);

hipError_t MTS_ENTRY(int flags, bool seq, bool f32, const LineArgs &a, hipStream_t s) {
  if (f32) {
    // the FP32 form: LOD and TS stages (Douglas-ADI runs in FP64 only)
    switch (flags) {
      MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump, float)
      MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF, float)
      MTS_LINE_CASE(kRhsCN | kFlagSolve | kFlagLhsJump, float)
      MTS_LINE_CASE(kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF, float)
      MTS_LINE_CASE(kRhsTSZ | kFlagSolve | kFlagLhsJump, float)
      case kRhsCN: return launch_flags<kRhsCN, float>(a.q.mmax, a, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (flags) {
    MTS_LINE_CASE(kRhsCopy | kFlagSolve, double)
    MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump, double)
    MTS_LINE_CASE(kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF, double)
    MTS_LINE_CASE(kRhsDouglas | kFlagSolve, double)
    MTS_LINE_CASE(kRhsCN | kFlagSolve | kFlagLhsJump, double)
    MTS_LINE_CASE(kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF, double)
    MTS_LINE_CASE(kRhsTSZ | kFlagSolve | kFlagLhsJump, double)
    case kRhsCN: return launch_flags<kRhsCN, double>(a.q.mmax, a, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace mts
