// gfx950 kernels of the Matched-TS per-step path.
//
// Every implicit stage of Douglas-ADI, LOD-IE/CN and Trapezoidal Splitting is
// a set of independent tridiagonal line systems (the reference solves each
// with TDMA, e.g. temporal/douglas_adi.cc:1784-1807).  Here one kernel
// launch does one stage along one direction for all lines:
//
//   * a workgroup owns a tile of 16 neighbouring lines, stages the stage's
//     input field(s) and the node classification for the whole tile in LDS
//     with coalesced loads (y/z lines: 16 consecutive x = one 128-byte row
//     segment per line row; x lines: contiguous rows),
//   * one 64-lane wavefront per line builds the right-hand side of its rows
//     (3-point beta-weighted stencil, analytic source and boundary values
//     from per-axis sin/cos tables, MIB fictitious-point rows from the sparse
//     block table), then solves the line with a partitioned Thomas / PCR
//     hybrid: each lane eliminates its own segment of rows in registers,
//     the 2-unknown-per-lane reduced system is folded to one unknown per lane
//     and solved by parallel cyclic reduction across the wavefront with lane
//     shuffles, and each lane back-substitutes its rows,
//   * the tile is written back with coalesced stores.
//
// All arithmetic mirrors the reference's operation order (fp contraction is
// disabled for this file) except the tridiagonal elimination order.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace mts {

namespace {

constexpr int kTile = 16;     // lines per workgroup tile
constexpr int kBlock = 256;   // threads per workgroup (4 wavefronts)

// ------------------------------------------------------------ geometry
__device__ __forceinline__ double tabv(const DevGeom &g, int a, int kind, int i) {
  return g.tab[a][kind * g.n[a] + i];
}
__device__ __forceinline__ size_t node_index(const DevGeom &g, int ix, int iy, int iz) {
  return ((size_t)iz * g.n[1] + iy) * (size_t)g.px + ix;
}

// ------------------------------------------------------------- physics
// u+ = e cos x cos y cos z (the boundary value and the outer exact solution)
__device__ __forceinline__ double outer_u(const DevGeom &g, double e, int ix, int iy, int iz) {
  return e * tabv(g, 0, kTabCos, ix) * tabv(g, 1, kTabCos, iy) * tabv(g, 2, kTabCos, iz);
}
// u- = e sin x sin y sin z
__device__ __forceinline__ double inner_u(const DevGeom &g, double e, int ix, int iy, int iz) {
  return e * tabv(g, 0, kTabSin, ix) * tabv(g, 1, kTabSin, iy) * tabv(g, 2, kTabSin, iz);
}

// f = (3 beta - 1) u - grad(beta).grad(u) on the node's side.
__device__ __forceinline__ double source(const PhysConst &P, const DevGeom &g, double e,
                                         bool outside, int ix, int iy, int iz) {
  if (outside) return P.f_out * outer_u(g, e, ix, iy, iz);
  double sx = tabv(g, 0, kTabSin, ix), sy = tabv(g, 1, kTabSin, iy), sz = tabv(g, 2, kTabSin, iz);
  double u = e * sx * sy * sz;
  if (P.beta_kind == 0) return P.f_in * u;
  double cx = tabv(g, 0, kTabCos, ix), cy = tabv(g, 1, kTabCos, iy);
  double b = P.b_in * (1.0 + 0.5 * sx * cy);
  double bx = 0.5 * P.b_in * cx * cy;
  double by = -0.5 * P.b_in * sx * sy;
  double ux = e * cx * sy * sz;
  double uy = e * sx * cy * sz;
  return (3.0 * b - 1.0) * u - bx * ux - by * uy;
}

// Operator weights beta(x -+ h/2)/h^2 along axis d on the node's side.
__device__ __forceinline__ void op_weights(const PhysConst &P, const DevGeom &g, int d,
                                           bool outside, int ix, int iy, int iz, double vd[3]) {
  if (outside || P.beta_kind == 0) {
    const int sd = outside ? 1 : 0;
    vd[0] = P.vdc[d][sd][0];
    vd[1] = P.vdc[d][sd][1];
    vd[2] = P.vdc[d][sd][2];
    return;
  }
  double b1, b2;
  if (d == 0) {
    double cy = tabv(g, 1, kTabCos, iy);
    b1 = P.b_in * (1.0 + 0.5 * tabv(g, 0, kTabSinM, ix) * cy);
    b2 = P.b_in * (1.0 + 0.5 * tabv(g, 0, kTabSinP, ix) * cy);
  } else if (d == 1) {
    double sx = tabv(g, 0, kTabSin, ix);
    b1 = P.b_in * (1.0 + 0.5 * sx * tabv(g, 1, kTabCosM, iy));
    b2 = P.b_in * (1.0 + 0.5 * sx * tabv(g, 1, kTabCosP, iy));
  } else {
    b1 = P.b_in * (1.0 + 0.5 * tabv(g, 0, kTabSin, ix) * tabv(g, 1, kTabCos, iy));
    b2 = b1;
  }
  vd[0] = b1 / P.hh[d];
  vd[1] = -(b1 + b2) / P.hh[d];
  vd[2] = b2 / P.hh[d];
}

// [u] and u_dir = P^{-1}[d][0] n.[beta grad u] at one crossing
// (Refresh_Jump, intersection.cc:2294-2490, approximate-jump mode).
__device__ __forceinline__ void jumps(const double *cr, double e, double &ju, double &jb) {
  const double sx = cr[0], sy = cr[1], sz = cr[2], cx = cr[3], cy = cr[4], cz = cr[5];
  const double bi = cr[6], bo = cr[7];
  const double me = -e;
  ju = e * cx * cy * cz - e * sx * sy * sz;
  double jbx = bo * (me * sx * cy * cz) - bi * (e * cx * sy * sz);
  double jby = bo * (me * cx * sy * cz) - bi * (e * sx * cy * sz);
  double jbz = bo * (me * cx * cy * sz) - bi * (e * sx * sy * cz);
  double xi = cr[8] * jbx + cr[9] * jby + cr[10] * jbz;
  jb = cr[11] * xi;
}

// corner jump combination over components 5..8 of a weight row
__device__ __forceinline__ double jl(const double *w, double juL, double jbL, double juR,
                                     double jbR) {
  return juL * w[5] + jbL * w[6] + juR * w[7] + jbR * w[8];
}

// ------------------------------------------- MIB fictitious-point stencils
// Sums of the patched rows without jump terms, in the reference's order
// (douglas_adi.cc:116-195).  F(i) returns the field value at line row i.
template <class F>
__device__ __forceinline__ double irr_sum0(const double *v, F f, int ll) {
  const double *vd = v + kIrrVd, *wr = v + kIrrWr;
  double s = vd[0] * f(ll - 1) + vd[1] * f(ll);
  for (int i = 0; i < 4; i++) s += vd[2] * f(ll - 1 + i) * wr[i];
  return s;
}
template <class F>
__device__ __forceinline__ double irr_sum1(const double *v, F f, int ll) {
  const double *vd = v + kIrrVd + 3, *wl = v + kIrrWl;
  double s = vd[1] * f(ll + 1) + vd[2] * f(ll + 2);
  for (int i = 0; i < 4; i++) s += vd[0] * f(ll - 1 + i) * wl[i];
  return s;
}
template <class F>
__device__ __forceinline__ double cor_sum0(const double *v, F f, int ll) {
  const double *vd = v + kCorVd, *w = v + kCorLwr;
  double s = vd[0] * f(ll - 1) + vd[1] * f(ll);
  for (int i = 0; i < 5; i++) s += vd[2] * f(ll - 1 + i) * w[i];
  return s;
}
template <class F>
__device__ __forceinline__ double cor_sum2(const double *v, F f, int ll) {
  const double *vd = v + kCorVd + 6, *w = v + kCorRwl;
  double s = vd[1] * f(ll + 2) + vd[2] * f(ll + 3);
  for (int i = 0; i < 5; i++) s += vd[0] * f(ll - 1 + i) * w[i];
  return s;
}
// middle corner row; with jumps interleaved as the Crank-Nicolson operators do
// (lod.cc:414-440) when cn is set, else the Douglas order.
template <class F>
__device__ __forceinline__ double cor_sum1(const double *v, F f, int ll, bool cn, double jL,
                                           double jR) {
  const double *vd = v + kCorVd + 3, *wl = v + kCorLwl, *wr = v + kCorRwr;
  double s = 0.0;
  for (int i = 0; i < 5; i++) s += vd[0] * f(ll - 1 + i) * wl[i];
  if (cn) s += vd[0] * jL;
  s += vd[1] * f(ll + 1);
  for (int i = 0; i < 5; i++) s += vd[2] * f(ll - 1 + i) * wr[i];
  if (cn) s += vd[2] * jR;
  return s;
}

// -------------------------------------------------- wavefront line solver
// Lane k holds rows [s_k, s_k + m) of one tridiagonal system
//   a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i
// (m >= 2 on active lanes, m == 0 on idle lanes; the first row of the system
// has a = 0 and the last c = 0).  On return the segment's first and last
// unknowns are in xs and xe and the rows in between in d (see solved_row).
template <int MMAX>
__device__ __forceinline__ void wave_solve(double (&a)[MMAX], double (&b)[MMAX],
                                           double (&c)[MMAX], double (&d)[MMAX], int m,
                                           int lane, double &xs, double &xe) {
  // reduced rows:  A1 x_{s-1} + B1 x_s + C1 x_e = D1   (first row)
  //                A2 x_s + B2 x_e + C2 x_{e+1} = D2   (last row)
  double A1 = 0.0, B1 = 1.0, C1 = 0.0, D1 = 0.0;
  double A2 = 0.0, B2 = 1.0, C2 = 0.0, D2 = 0.0;
  if (m >= 2) {
    // downward elimination from row 1 keeping x_s as a parameter:
    // alpha_i x_s + x_i + gamma_i x_{i+1} = delta_i  (stored in a, c, d)
    {
      double r = 1.0 / b[1];
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
        double r = 1.0 / (b[i] - a[i] * c[i - 1]);
        d[i] = (d[i] - a[i] * d[i - 1]) * r;
        a[i] = -a[i] * a[i - 1] * r;
        c[i] = c[i] * r;
        // the last row reached is the segment's last row
        A2 = a[i];
        C2 = c[i];
        D2 = d[i];
      }
    }
    // upward elimination down to row 1, expressing x_i through x_s and x_e
#pragma unroll
    for (int i = MMAX - 3; i >= 1; i--) {
      if (i <= m - 3) {
        d[i] = d[i] - c[i] * d[i + 1];
        a[i] = a[i] - c[i] * a[i + 1];
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
      B1 = b[0] - c[0] * a[1];
      C1 = -c[0] * c[1];
      D1 = d[0] - c[0] * d[1];
    }
  }
  // fold x_s out: one equation per lane in the segment-end unknowns x_e
  double A1n = __shfl_down(A1, 1, 64);
  double B1n = __shfl_down(B1, 1, 64);
  double C1n = __shfl_down(C1, 1, 64);
  double D1n = __shfl_down(D1, 1, 64);
  double ea = 0.0, eb = 1.0, ec = 0.0, ed = 0.0;
  if (m >= 2) {
    if (lane == 63) {
      A1n = 0.0;
      B1n = 1.0;
      C1n = 0.0;
      D1n = 0.0;
    }
    double i1 = 1.0 / B1, in = 1.0 / B1n;
    ea = -A2 * A1 * i1;
    eb = B2 - A2 * C1 * i1 - C2 * A1n * in;
    ec = -C2 * C1n * in;
    ed = D2 - A2 * D1 * i1 - C2 * D1n * in;
  }
  // parallel cyclic reduction across the wavefront
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    double am = __shfl_up(ea, s, 64), bm = __shfl_up(eb, s, 64);
    double cm = __shfl_up(ec, s, 64), dm = __shfl_up(ed, s, 64);
    double ap = __shfl_down(ea, s, 64), bp = __shfl_down(eb, s, 64);
    double cp = __shfl_down(ec, s, 64), dp = __shfl_down(ed, s, 64);
    if (lane < s) {
      am = 0.0; bm = 1.0; cm = 0.0; dm = 0.0;
    }
    if (lane + s >= 64) {
      ap = 0.0; bp = 1.0; cp = 0.0; dp = 0.0;
    }
    double k1 = ea / bm, k2 = ec / bp;
    double na = -am * k1;
    double nc = -cp * k2;
    double nb = eb - cm * k1 - ap * k2;
    double nd = ed - dm * k1 - dp * k2;
    ea = na;
    eb = nb;
    ec = nc;
    ed = nd;
  }
  xe = ed / eb;
  double xprev = __shfl_up(xe, 1, 64);
  if (lane == 0) xprev = 0.0;
  xs = xe;
  if (m >= 2) {
    xs = (D1 - A1 * xprev - C1 * xe) / B1;
#pragma unroll
    for (int i = 1; i < MMAX - 1; i++) {
      if (i < m - 1) d[i] = d[i] - a[i] * xs - c[i] * xe;
    }
  }
}

// value of local row r after wave_solve
template <int MMAX>
__device__ __forceinline__ double solved_row(const double (&d)[MMAX], int r, int m, double xs,
                                             double xe) {
  return r == 0 ? xs : (r == m - 1 ? xe : d[r]);
}

// segment of lane for a line of n rows: p lanes active, rows balanced
__device__ __forceinline__ void segment(int n, int lane, int &s, int &m) {
  int p = n / 2;
  if (p > 64) p = 64;
  int q = n / p, rem = n - q * p;
  if (lane < p) {
    m = q + (lane < rem ? 1 : 0);
    s = lane * q + (lane < rem ? lane : rem);
  } else {
    m = 0;
    s = n;
  }
}

// ----------------------------------------------------------- line kernel
template <int DIR>
struct LineMap {
  // number of lines along the tile index j and the outer index k
  __device__ static int nj(const DevGeom &g) { return DIR == 0 ? g.n[1] : g.n[0]; }
  __device__ static int nk(const DevGeom &g) { return DIR == 2 ? g.n[1] : g.n[2]; }
  __device__ static size_t stride(const DevGeom &g) {
    return DIR == 0 ? 1 : DIR == 1 ? (size_t)g.px : (size_t)g.n[1] * g.px;
  }
  __device__ static size_t base(const DevGeom &g, int j, int k) {
    return DIR == 0 ? ((size_t)k * g.n[1] + j) * g.px
                    : DIR == 1 ? (size_t)k * g.n[1] * g.px + j : (size_t)k * g.px + j;
  }
  __device__ static void node(int j, int k, int i, int &ix, int &iy, int &iz) {
    if (DIR == 0) { ix = i; iy = j; iz = k; }
    else if (DIR == 1) { ix = j; iy = i; iz = k; }
    else { ix = j; iy = k; iz = i; }
  }
};

template <int DIR, int MMAX, int FLAGS>
__global__ __launch_bounds__(kBlock) void line_kernel(LineArgs p) {
  constexpr int RHS = FLAGS & 3;
  constexpr bool LJ = (FLAGS & kFlagLhsJump) != 0;
  constexpr bool SOLVE = (FLAGS & kFlagSolve) != 0;
  constexpr bool EPI = (FLAGS & kFlagEpiF) != 0;
  constexpr bool HAS_S = (RHS == kRhsDouglas);
  constexpr bool TWO = HAS_S || !SOLVE;  // second LDS tile: S, or the output of explicit stages
  using M = LineMap<DIR>;
  const DevGeom &g = p.g;
  const int n = g.n[DIR];
  const int ns = n + 1;
  const int nj = M::nj(g);
  const int j0 = blockIdx.x * kTile;
  const int k = blockIdx.y + 1;
  const size_t rs = M::stride(g);

  extern __shared__ double smem[];
  double *tA = smem;
  double *tB = smem + kTile * ns;
  int8_t *tI = reinterpret_cast<int8_t *>(smem + (TWO ? 2 : 1) * kTile * ns);

  // ---- stage the tile
  for (int e = threadIdx.x; e < kTile * n; e += blockDim.x) {
    int l, r;
    if (DIR == 0) {
      l = e / n;
      r = e - l * n;
    } else {
      l = e & (kTile - 1);
      r = e / kTile;
    }
    int j = j0 + l;
    if (j >= nj) continue;
    size_t gi = M::base(g, j, k) + (size_t)r * rs;
    double va = p.A[gi];
    if (p.in_analytic && (r == 0 || r == n - 1)) {
      int ix, iy, iz;
      M::node(j, k, r, ix, iy, iz);
      va = outer_u(g, p.e_in, ix, iy, iz);
    }
    tA[l * ns + r] = va;
    if (HAS_S) tB[l * ns + r] = p.S[gi];
    tI[l * ns + r] = g.ind[gi];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwave = blockDim.x >> 6;
  int s, m;
  segment(n, lane, s, m);
  const double tdt = p.theta_dt;
  const double mtdt = -tdt;

  for (int l = wave; l < kTile; l += nwave) {
    const int j = j0 + l;
    if (j < 1 || j > nj - 2) continue;
    const double *la = tA + l * ns;
    const double *lb = tB + l * ns;
    const int8_t *li = tI + l * ns;
    auto FA = [&](int i) { return la[i]; };
    auto FS = [&](int i) { return lb[i]; };

    double a[MMAX], b[MMAX], c[MMAX], d[MMAX];
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      a[r] = 0.0;
      b[r] = 1.0;
      c[r] = 0.0;
      d[r] = 0.0;
      if (r < m) {
        const int i = s + r;
        int ix, iy, iz;
        M::node(j, k, i, ix, iy, iz);
        if (i == 0 || i == n - 1) {
          d[r] = outer_u(g, p.e_out, ix, iy, iz);
        } else {
          const bool out = li[i] > 0;
          double vd[3];
          op_weights(p.P, g, DIR, out, ix, iy, iz, vd);
          a[r] = mtdt * vd[0];
          b[r] = 1 - tdt * vd[1];
          c[r] = mtdt * vd[2];
          if (RHS == kRhsCopy) {
            d[r] = la[i];
          } else if (RHS == kRhsDouglas) {
            double sum = lb[i - 1] * vd[0] + lb[i] * vd[1] + lb[i + 1] * vd[2];
            d[r] = -p.dt * sum + la[i];
          } else {
            double sum = la[i - 1] * vd[0] + la[i] * vd[1] + la[i + 1] * vd[2];
            double v = tdt * sum + la[i];
            if (RHS == kRhsTSZ) {
              v = v + p.c_f1 * source(p.P, g, p.e_f1, out, ix, iy, iz);
              v = v + p.c_f2 * source(p.P, g, p.e_f2, out, ix, iy, iz);
            }
            d[r] = v;
          }
        }
      }
    }

    // ---- MIB blocks of this line
    const int lid = k * nj + j;
    const int bb = p.mib.blk_ptr[lid], be = p.mib.blk_ptr[lid + 1];
    for (int bi = bb; bi < be; bi++) {
      const BlockHdr h = p.mib.hdr[bi];
      const int ll = h.ll;
      const int nr = (h.type == kBlkIrr) ? 2 : 3;
      if (m == 0 || ll + nr - 1 < s || ll >= s + m) continue;
      const double *v = p.mib.val + h.off;
      double ju = 0.0, jb = 0.0, juR = 0.0, jbR = 0.0;
      if (RHS == kRhsCN || RHS == kRhsTSZ || LJ) {
        jumps(p.mib.cr + (size_t)h.c0 * kCrLen, p.e_jump, ju, jb);
        if (h.type == kBlkCor) jumps(p.mib.cr + (size_t)(h.c0 + 1) * kCrLen, p.e_jump, juR, jbR);
      }
      double bv[3];
      const double *lhs;
      if (h.type == kBlkIrr) {
        const double *vd0 = v + kIrrVd, *vd1 = v + kIrrVd + 3;
        const double *wl = v + kIrrWl, *wr = v + kIrrWr;
        const double J0 = ju * wr[4] + jb * wr[5];
        const double J1 = ju * wl[4] + jb * wl[5];
        if (RHS == kRhsCopy) {
          bv[0] = la[ll];
          bv[1] = la[ll + 1];
        } else if (RHS == kRhsDouglas) {
          bv[0] = -p.dt * irr_sum0(v, FS, ll) + la[ll];
          bv[1] = -p.dt * irr_sum1(v, FS, ll) + la[ll + 1];
        } else {
          double s0 = irr_sum0(v, FA, ll) + vd0[2] * J0;
          double s1 = irr_sum1(v, FA, ll) + vd1[0] * J1;
          bv[0] = tdt * s0 + la[ll];
          bv[1] = tdt * s1 + la[ll + 1];
        }
        if (RHS == kRhsTSZ) {
          for (int r = 0; r < 2; r++) {
            int ix, iy, iz;
            M::node(j, k, ll + r, ix, iy, iz);
            const bool out = li[ll + r] > 0;
            bv[r] = bv[r] + p.c_f1 * source(p.P, g, p.e_f1, out, ix, iy, iz);
            bv[r] = bv[r] + p.c_f2 * source(p.P, g, p.e_f2, out, ix, iy, iz);
          }
        }
        if (LJ) {
          bv[0] = bv[0] + tdt * vd0[2] * J0;
          bv[1] = bv[1] + tdt * vd1[0] * J1;
        }
        const double *mu = v + kIrrMul;
        bv[1] -= mu[0] * bv[0];
        bv[0] -= mu[1] * bv[1];
        lhs = v + kIrrLhs;
      } else {
        const double *vd0 = v + kCorVd, *vd1 = v + kCorVd + 3, *vd2 = v + kCorVd + 6;
        const double JLwr = jl(v + kCorLwr, ju, jb, juR, jbR);
        const double JLwl = jl(v + kCorLwl, ju, jb, juR, jbR);
        const double JRwr = jl(v + kCorRwr, ju, jb, juR, jbR);
        const double JRwl = jl(v + kCorRwl, ju, jb, juR, jbR);
        if (RHS == kRhsCopy) {
          bv[0] = la[ll];
          bv[1] = la[ll + 1];
          bv[2] = la[ll + 2];
        } else if (RHS == kRhsDouglas) {
          bv[0] = -p.dt * cor_sum0(v, FS, ll) + la[ll];
          bv[1] = -p.dt * cor_sum1(v, FS, ll, false, 0.0, 0.0) + la[ll + 1];
          bv[2] = -p.dt * cor_sum2(v, FS, ll) + la[ll + 2];
        } else {
          double s0 = cor_sum0(v, FA, ll) + vd0[2] * JLwr;
          double s1 = cor_sum1(v, FA, ll, true, JLwl, JRwr);
          double s2 = cor_sum2(v, FA, ll) + vd2[0] * JRwl;
          bv[0] = tdt * s0 + la[ll];
          bv[1] = tdt * s1 + la[ll + 1];
          bv[2] = tdt * s2 + la[ll + 2];
        }
        if (RHS == kRhsTSZ) {
          for (int r = 0; r < 3; r++) {
            int ix, iy, iz;
            M::node(j, k, ll + r, ix, iy, iz);
            const bool out = li[ll + r] > 0;
            bv[r] = bv[r] + p.c_f1 * source(p.P, g, p.e_f1, out, ix, iy, iz);
            bv[r] = bv[r] + p.c_f2 * source(p.P, g, p.e_f2, out, ix, iy, iz);
          }
        }
        if (LJ) {
          bv[0] = bv[0] + tdt * vd0[2] * JLwr;
          bv[1] = bv[1] + tdt * vd1[0] * JLwl;
          bv[1] += tdt * vd1[2] * JRwr;
          bv[2] = bv[2] + tdt * vd2[0] * JRwl;
        }
        const double *mu = v + kCorMul;
        bv[1] -= mu[0] * bv[0];
        bv[2] -= mu[1] * bv[0];
        bv[0] -= mu[2] * bv[2];
        bv[1] -= mu[3] * bv[2];
        bv[2] -= mu[4] * bv[1];
        bv[0] -= mu[5] * bv[1];
        lhs = v + kCorLhs;
      }
#pragma unroll
      for (int r = 0; r < MMAX; r++) {
        const int rr = s + r - ll;
        if (r < m && rr >= 0 && rr < nr) {
          if (SOLVE) {
            a[r] = lhs[rr * 3 + 0];
            b[r] = lhs[rr * 3 + 1];
            c[r] = lhs[rr * 3 + 2];
          }
          d[r] = rr == 0 ? bv[0] : rr == 1 ? bv[1] : bv[2];
        }
      }
    }

    double xs = 0.0, xe = 0.0;
    if (SOLVE) wave_solve<MMAX>(a, b, c, d, m, lane, xs, xe);

    double *lo = TWO && !HAS_S ? tB + l * ns : tA + l * ns;
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      if (r < m) {
        const int i = s + r;
        double x = SOLVE ? solved_row(d, r, m, xs, xe) : d[r];
        if (EPI) {
          int ix, iy, iz;
          M::node(j, k, i, ix, iy, iz);
          x = x + p.c_f1 * source(p.P, g, p.e_f1, li[i] > 0, ix, iy, iz);
        }
        lo[i] = x;
      }
    }
  }
  __syncthreads();

  // ---- write the tile's interior lines back
  const double *tO = TWO && !HAS_S ? tB : tA;
  for (int e = threadIdx.x; e < kTile * n; e += blockDim.x) {
    int l, r;
    if (DIR == 0) {
      l = e / n;
      r = e - l * n;
    } else {
      l = e & (kTile - 1);
      r = e / kTile;
    }
    int j = j0 + l;
    if (j < 1 || j > nj - 2) continue;
    p.out[M::base(g, j, k) + (size_t)r * rs] = tO[l * ns + r];
  }
}

template <int DIR, int MMAX, int FLAGS>
hipError_t launch_line_t(const LineArgs &a, hipStream_t st) {
  using M = LineMap<DIR>;
  const DevGeom &g = a.g;
  const int n = g.n[DIR];
  const int nj = DIR == 0 ? g.n[1] : g.n[0];
  const int nk = DIR == 2 ? g.n[1] : g.n[2];
  (void)sizeof(M);
  constexpr bool TWO = ((FLAGS & 3) == kRhsDouglas) || !(FLAGS & kFlagSolve);
  size_t shm = (size_t)kTile * (n + 1) * sizeof(double) * (TWO ? 2 : 1) + (size_t)kTile * (n + 1);
  dim3 grid((nj + kTile - 1) / kTile, nk - 2);
  hipLaunchKernelGGL((line_kernel<DIR, MMAX, FLAGS>), grid, dim3(kBlock), shm, st, a);
  return hipGetLastError();
}

template <int MMAX, int FLAGS>
hipError_t launch_line_m(int dir, const LineArgs &a, hipStream_t st) {
  if (dir == 0) return launch_line_t<0, MMAX, FLAGS>(a, st);
  if (dir == 1) return launch_line_t<1, MMAX, FLAGS>(a, st);
  return launch_line_t<2, MMAX, FLAGS>(a, st);
}

template <int FLAGS>
hipError_t launch_line_f(int dir, const LineArgs &a, hipStream_t st) {
  const int n = a.g.n[dir];
  int p = n / 2;
  if (p > 64) p = 64;
  const int mmax = (n + p - 1) / p;
  if (mmax <= 4) return launch_line_m<4, FLAGS>(dir, a, st);
  if (mmax <= 8) return launch_line_m<8, FLAGS>(dir, a, st);
  if (mmax <= 16) return launch_line_m<16, FLAGS>(dir, a, st);
  return hipErrorInvalidValue;
}

// ------------------------------------------------ Douglas stage-1 RHS
__global__ __launch_bounds__(256) void rhs1_kernel(Rhs1Args p) {
  const DevGeom &g = p.g;
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y;
  const int iz = blockIdx.z;
  if (ix >= g.n[0]) return;
  const size_t c = node_index(g, ix, iy, iz);
  if (ix == 0 || iy == 0 || iz == 0 || ix == g.n[0] - 1 || iy == g.n[1] - 1 ||
      iz == g.n[2] - 1) {
    p.R[c] = outer_u(g, p.e, ix, iy, iz);
    return;
  }
  const bool out = g.ind[c] > 0;
  const size_t sy = (size_t)g.px, sz = (size_t)g.n[1] * g.px;
  double vy[3], vz[3];
  op_weights(p.P, g, 1, out, ix, iy, iz, vy);
  op_weights(p.P, g, 2, out, ix, iy, iz, vz);
  const double u = p.u[c];
  double sum_y = p.u[c - sy] * vy[0] + u * vy[1] + p.u[c + sy] * vy[2];
  double sum_z = p.u[c - sz] * vz[0] + u * vz[1] + p.u[c + sz] * vz[2];
  double r = p.dt * sum_y;
  r += p.dt * sum_z;
  r += u;
  r += p.dt * source(p.P, g, p.e, out, ix, iy, iz);
  p.R[c] = r;
}

// src += jump terms of block row rr (Src_2nd, douglas_adi.cc:1123-1428)
__device__ __forceinline__ double src_jump(const DevMib &mib, int bidx, int rr, double e,
                                           double src) {
  const BlockHdr h = mib.hdr[bidx];
  const double *v = mib.val + h.off;
  double ju, jb;
  jumps(mib.cr + (size_t)h.c0 * kCrLen, e, ju, jb);
  if (h.type == kBlkIrr) {
    if (rr == 0) {
      const double *w = v + kIrrWr;
      src += v[kIrrVd + 2] * (ju * w[4] + jb * w[5]);
    } else {
      const double *w = v + kIrrWl;
      src += v[kIrrVd + 3] * (ju * w[4] + jb * w[5]);
    }
    return src;
  }
  double juR, jbR;
  jumps(mib.cr + (size_t)(h.c0 + 1) * kCrLen, e, juR, jbR);
  if (rr == 0) {
    src += v[kCorVd + 2] * jl(v + kCorLwr, ju, jb, juR, jbR);
  } else if (rr == 1) {
    src += v[kCorVd + 3] * jl(v + kCorLwl, ju, jb, juR, jbR);
    src += v[kCorVd + 5] * jl(v + kCorRwr, ju, jb, juR, jbR);
  } else {
    src += v[kCorVd + 6] * jl(v + kCorRwl, ju, jb, juR, jbR);
  }
  return src;
}

// fictitious-point stencil sum at block row rr, reading the field from global memory
__device__ __forceinline__ double block_sum(const DevMib &mib, int bidx, int rr,
                                            const double *u, size_t base, size_t rs) {
  const BlockHdr h = mib.hdr[bidx];
  const double *v = mib.val + h.off;
  auto F = [&](int i) { return u[base + (size_t)i * rs]; };
  if (h.type == kBlkIrr) return rr == 0 ? irr_sum0(v, F, h.ll) : irr_sum1(v, F, h.ll);
  if (rr == 0) return cor_sum0(v, F, h.ll);
  if (rr == 1) return cor_sum1(v, F, h.ll, false, 0.0, 0.0);
  return cor_sum2(v, F, h.ll);
}

__global__ __launch_bounds__(256) void spec_kernel(SpecArgs p) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.nspec) return;
  const DevGeom &g = p.g;
  const int *rec = p.spec + (size_t)t * kSpecLen;
  const int ix = rec[0], iy = rec[1], iz = rec[2];
  const size_t c = node_index(g, ix, iy, iz);
  const bool out = g.ind[c] > 0;
  const size_t sy = (size_t)g.px, sz = (size_t)g.n[1] * g.px;
  const double u = p.u[c];
  double sum_y, sum_z;
  if (rec[3 + 1] >= 0) {
    sum_y = block_sum(p.mib[1], rec[3 + 1], rec[6 + 1], p.u, c - (size_t)iy * sy, sy);
  } else {
    double vy[3];
    op_weights(p.P, g, 1, out, ix, iy, iz, vy);
    sum_y = p.u[c - sy] * vy[0] + u * vy[1] + p.u[c + sy] * vy[2];
  }
  if (rec[3 + 2] >= 0) {
    sum_z = block_sum(p.mib[2], rec[3 + 2], rec[6 + 2], p.u, c - (size_t)iz * sz, sz);
  } else {
    double vz[3];
    op_weights(p.P, g, 2, out, ix, iy, iz, vz);
    sum_z = p.u[c - sz] * vz[0] + u * vz[1] + p.u[c + sz] * vz[2];
  }
  double src = source(p.P, g, p.e, out, ix, iy, iz);
  for (int d = 0; d < 3; d++)
    if (rec[3 + d] >= 0) src = src_jump(p.mib[d], rec[3 + d], rec[6 + d], p.e, src);
  double r = p.dt * sum_y;
  r += p.dt * sum_z;
  r += u;
  r += p.dt * src;
  p.R[c] = r;
}

// ------------------------------------------------------- boundary faces
__global__ __launch_bounds__(256) void faces_kernel(DevGeom g, PhysConst P, double *out,
                                                    double e, double cf) {
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const long long fz = (long long)nx * ny, fy = (long long)nx * nz, fx = (long long)ny * nz;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int ix, iy, iz;
  if (t < 2 * fz) {
    iz = t < fz ? 0 : nz - 1;
    long long q = t % fz;
    ix = (int)(q % nx);
    iy = (int)(q / nx);
  } else if ((t -= 2 * fz) < 2 * fy) {
    iy = t < fy ? 0 : ny - 1;
    long long q = t % fy;
    ix = (int)(q % nx);
    iz = (int)(q / nx);
  } else if ((t -= 2 * fy) < 2 * fx) {
    ix = t < fx ? 0 : nx - 1;
    long long q = t % fx;
    iy = (int)(q % ny);
    iz = (int)(q / ny);
  } else {
    return;
  }
  const size_t c = node_index(g, ix, iy, iz);
  double v = outer_u(g, e, ix, iy, iz);
  if (cf != 0.0) v = v + cf * source(P, g, e, g.ind[c] > 0, ix, iy, iz);
  out[c] = v;
}

// ------------------------------------------------------- initialization
__global__ __launch_bounds__(256) void init_kernel(DevGeom g, double *u, double e) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y, iz = blockIdx.z;
  if (ix >= g.n[0]) return;
  const size_t c = node_index(g, ix, iy, iz);
  u[c] = g.ind[c] > 0 ? outer_u(g, e, ix, iy, iz) : inner_u(g, e, ix, iy, iz);
}

// ------------------------------------------------------------ error norms
// Reference order index (ix outer, iz inner, douglas_adi.cc:1736-1760) so the
// first maximum in that order wins ties.
__global__ __launch_bounds__(256) void norms_kernel(DevGeom g, const double *u, double e,
                                                    double *pmax, long long *pidx,
                                                    double *psq) {
  __shared__ double smax[256];
  __shared__ long long sidx[256];
  __shared__ double ssq[256];
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const long long total = (long long)nx * ny * nz;
  double mx = -1.0, sq = 0.0;
  long long mi = total;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int ix = (int)(t % nx);
    long long r = t / nx;
    int iy = (int)(r % ny);
    int iz = (int)(r / ny);
    size_t c = node_index(g, ix, iy, iz);
    double ex = g.ind[c] > 0 ? outer_u(g, e, ix, iy, iz) : inner_u(g, e, ix, iy, iz);
    double er = fabs(u[c] - ex);
    long long ri = ((long long)ix * ny + iy) * nz + iz;
    if (er > mx || (er == mx && ri < mi)) {
      mx = er;
      mi = ri;
    }
    sq += er * er;
  }
  smax[threadIdx.x] = mx;
  sidx[threadIdx.x] = mi;
  ssq[threadIdx.x] = sq;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double om = smax[threadIdx.x + w];
      long long oi = sidx[threadIdx.x + w];
      if (om > smax[threadIdx.x] || (om == smax[threadIdx.x] && oi < sidx[threadIdx.x])) {
        smax[threadIdx.x] = om;
        sidx[threadIdx.x] = oi;
      }
      ssq[threadIdx.x] += ssq[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmax[blockIdx.x] = smax[0];
    pidx[blockIdx.x] = sidx[0];
    psq[blockIdx.x] = ssq[0];
  }
}

// ----------------------------------------------------- batched tridiagonal
template <int MMAX>
__global__ __launch_bounds__(256) void tdma_kernel(int n, int nlines, const double *A,
                                                   const double *B, const double *C,
                                                   const double *R, double *U) {
  const int lane = threadIdx.x & 63;
  const int line = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (line >= nlines) return;  // uniform per wavefront
  int s, m;
  segment(n, lane, s, m);
  const size_t base = (size_t)line * n;
  double a[MMAX], b[MMAX], c[MMAX], d[MMAX];
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    a[r] = 0.0;
    b[r] = 1.0;
    c[r] = 0.0;
    d[r] = 0.0;
    if (r < m) {
      const int i = s + r;
      a[r] = i == 0 ? 0.0 : A[base + i];
      b[r] = B[base + i];
      c[r] = i == n - 1 ? 0.0 : C[base + i];
      d[r] = R[base + i];
    }
  }
  double xs, xe;
  wave_solve<MMAX>(a, b, c, d, m, lane, xs, xe);
#pragma unroll
  for (int r = 0; r < MMAX; r++)
    if (r < m) U[base + s + r] = solved_row(d, r, m, xs, xe);
}

}  // namespace

// --------------------------------------------------------------- launchers
hipError_t launch_line(int dir, int flags, const LineArgs &a, hipStream_t s) {
  switch (flags) {
    case kRhsCopy | kFlagSolve: return launch_line_f<kRhsCopy | kFlagSolve>(dir, a, s);
    case kRhsCopy | kFlagSolve | kFlagLhsJump:
      return launch_line_f<kRhsCopy | kFlagSolve | kFlagLhsJump>(dir, a, s);
    case kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF:
      return launch_line_f<kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF>(dir, a, s);
    case kRhsDouglas | kFlagSolve: return launch_line_f<kRhsDouglas | kFlagSolve>(dir, a, s);
    case kRhsCN | kFlagSolve | kFlagLhsJump:
      return launch_line_f<kRhsCN | kFlagSolve | kFlagLhsJump>(dir, a, s);
    case kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF:
      return launch_line_f<kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF>(dir, a, s);
    case kRhsCN: return launch_line_f<kRhsCN>(dir, a, s);
    case kRhsTSZ | kFlagSolve | kFlagLhsJump:
      return launch_line_f<kRhsTSZ | kFlagSolve | kFlagLhsJump>(dir, a, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_rhs1(const Rhs1Args &a, hipStream_t s) {
  dim3 grid((a.g.n[0] + 255) / 256, a.g.n[1], a.g.n[2]);
  hipLaunchKernelGGL(rhs1_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_spec(const SpecArgs &a, hipStream_t s) {
  if (a.nspec == 0) return hipSuccess;
  hipLaunchKernelGGL(spec_kernel, dim3((a.nspec + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_faces(const DevGeom &g, const PhysConst &P, double *out, double e, double cf,
                        hipStream_t s) {
  long long total = 2LL * ((long long)g.n[0] * g.n[1] + (long long)g.n[0] * g.n[2] +
                           (long long)g.n[1] * g.n[2]);
  hipLaunchKernelGGL(faces_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g, P,
                     out, e, cf);
  return hipGetLastError();
}

hipError_t launch_init(const DevGeom &g, double *u, double e, hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  hipLaunchKernelGGL(init_kernel, grid, dim3(256), 0, s, g, u, e);
  return hipGetLastError();
}

int norms_blocks(const DevGeom &g) {
  long long total = (long long)g.n[0] * g.n[1] * g.n[2];
  long long b = (total + 255) / 256;
  return (int)(b < 2048 ? b : 2048);
}

hipError_t launch_norms(const DevGeom &g, const double *u, double e, double *part_max,
                        long long *part_idx, double *part_sq, int *nparts, hipStream_t s) {
  int nb = norms_blocks(g);
  *nparts = nb;
  hipLaunchKernelGGL(norms_kernel, dim3(nb), dim3(256), 0, s, g, u, e, part_max, part_idx,
                     part_sq);
  return hipGetLastError();
}

hipError_t launch_tdma(int n, int nlines, const double *a, const double *b, const double *c,
                       const double *r, double *u, hipStream_t s) {
  if (n < 2 || nlines < 1) return hipErrorInvalidValue;
  int p = n / 2;
  if (p > 64) p = 64;
  const int mmax = (n + p - 1) / p;
  dim3 grid((nlines + 3) / 4);
  if (mmax <= 4) hipLaunchKernelGGL((tdma_kernel<4>), grid, dim3(256), 0, s, n, nlines, a, b, c, r, u);
  else if (mmax <= 8) hipLaunchKernelGGL((tdma_kernel<8>), grid, dim3(256), 0, s, n, nlines, a, b, c, r, u);
  else if (mmax <= 16) hipLaunchKernelGGL((tdma_kernel<16>), grid, dim3(256), 0, s, n, nlines, a, b, c, r, u);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace mts
