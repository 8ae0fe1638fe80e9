// Per-node device functions of the per-step path, compiled as HIP device
// code by kernels.hip and as plain host C++ by the CPU bounds-check build
// (cpu_check.cpp), so every index a kernel forms can be checked on the host.
#pragma once

#include <cmath>
#include <cstddef>

#include "device_types.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MTS_HD __host__ __device__ __forceinline__
#else
#define MTS_HD inline
#endif
// Keeps the instruction scheduler from hoisting the loads of later phases of
// an MIB block above earlier ones (all 60 doubles of a corner record live at
// once would cap the kernel's occupancy).
#if defined(__HIP_DEVICE_COMPILE__)
#define MTS_PHASE() __builtin_amdgcn_sched_barrier(0)
#else
#define MTS_PHASE() ((void)0)
#endif
#if defined(__HIPCC__)
#define MTS_OUTLINE __host__ __device__ __attribute__((noinline))
#else
#define MTS_OUTLINE inline
#endif

namespace mts {

// a * b + c rounded once, in the arithmetic of the arguments
MTS_HD double fmad(double a, double b, double c) { return fma(a, b, c); }
MTS_HD float fmad(float a, float b, float c) { return fmaf(a, b, c); }

// Two neighbouring doubles at a 16-byte aligned address, as one load on the
// device.
MTS_HD void ld2(const double *a, double (&v)[2]) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double2 t = *reinterpret_cast<const double2 *>(a);
  v[0] = t.x;
  v[1] = t.y;
#else
  v[0] = a[0];
  v[1] = a[1];
#endif
}
// The same for two floats at an 8-byte aligned address (the FP32 field),
// widened.
MTS_HD void ld2(const float *a, double (&v)[2]) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float2 t = *reinterpret_cast<const float2 *>(a);
  v[0] = (double)t.x;
  v[1] = (double)t.y;
#else
  v[0] = (double)a[0];
  v[1] = (double)a[1];
#endif
}

// ------------------------------------------------------------ geometry
MTS_HD double tabv(const DevGeom &g, int a, int kind, int i) {
  return g.tab[a][kind * g.n[a] + i];
}
MTS_HD size_t node_index(const DevGeom &g, int ix, int iy, int iz) {
  return ((size_t)iz * g.n[1] + iy) * (size_t)g.px + ix;
}

// ------------------------------------------------------------- physics
// The per-node physics below reads the trigonometric tables through an
// accessor tv(a, kind) (the entry of axis a at the node), so that one text of
// each formula serves both callers that load entries where they are used and
// callers that load them all up front.
template <class TV>
MTS_HD double outer_with(double e, const TV &tv) {
  return e * tv(0, kTabCos) * tv(1, kTabCos) * tv(2, kTabCos);
}
// f = (3 beta - 1) u - grad(beta).grad(u) on the node's side.
template <class TV>
MTS_HD double source_with(const PhysConst &P, double e, bool outside, const TV &tv) {
  if (outside) return P.f_out * outer_with(e, tv);
  double sx = tv(0, kTabSin), sy = tv(1, kTabSin), sz = tv(2, kTabSin);
  double u = e * sx * sy * sz;
  if (P.beta_kind == 0) return P.f_in * u;
  double cx = tv(0, kTabCos), cy = tv(1, kTabCos);
  double b = P.b_in * (1.0 + 0.5 * sx * cy);
  double bx = 0.5 * P.b_in * cx * cy;
  double by = -0.5 * P.b_in * sx * sy;
  double ux = e * cx * sy * sz;
  double uy = e * sx * cy * sz;
  return (3.0 * b - 1.0) * u - bx * ux - by * uy;
}
// Operator weights beta(x -+ h/2)/h^2 along axis d on the node's side.
template <class TV>
MTS_HD void op_weights_with(const PhysConst &P, int d, bool outside, const TV &tv, double vd[3]) {
  if (outside || P.beta_kind == 0) {
    const int sd = outside ? 1 : 0;
    vd[0] = P.vdc[d][sd][0];
    vd[1] = P.vdc[d][sd][1];
    vd[2] = P.vdc[d][sd][2];
    return;
  }
  double b1, b2;
  if (d == 0) {
    double cy = tv(1, kTabCos);
    b1 = P.b_in * (1.0 + 0.5 * tv(0, kTabSinM) * cy);
    b2 = P.b_in * (1.0 + 0.5 * tv(0, kTabSinP) * cy);
  } else if (d == 1) {
    double sx = tv(0, kTabSin);
    b1 = P.b_in * (1.0 + 0.5 * sx * tv(1, kTabCosM));
    b2 = P.b_in * (1.0 + 0.5 * sx * tv(1, kTabCosP));
  } else {
    b1 = P.b_in * (1.0 + 0.5 * tv(0, kTabSin) * tv(1, kTabCos));
    b2 = b1;
  }
  vd[0] = b1 * P.ihh[d];
  vd[1] = -(b1 + b2) * P.ihh[d];
  vd[2] = b2 * P.ihh[d];
}
// The accessor that loads each entry where it is used.
struct TabAt {
  const DevGeom &g;
  int i[3];
  MTS_HD double operator()(int a, int kind) const { return tabv(g, a, kind, i[a]); }
};

// u+ = e cos x cos y cos z (the boundary value and the outer exact solution)
MTS_HD double outer_u(const DevGeom &g, double e, int ix, int iy, int iz) {
  return outer_with(e, TabAt{g, {ix, iy, iz}});
}
// u- = e sin x sin y sin z
MTS_HD double inner_u(const DevGeom &g, double e, int ix, int iy, int iz) {
  return e * tabv(g, 0, kTabSin, ix) * tabv(g, 1, kTabSin, iy) * tabv(g, 2, kTabSin, iz);
}

MTS_HD double source(const PhysConst &P, const DevGeom &g, double e,
                                         bool outside, int ix, int iy, int iz) {
  return source_with(P, e, outside, TabAt{g, {ix, iy, iz}});
}

// Value of boundary node (ix, iy, iz): the exact outer solution at e, plus
// cf f(e) when cf != 0 (LOD's source epilogue reaches the boundary too).
MTS_HD double face_value(const PhysConst &P, const DevGeom &g, double e, double cf, int ix,
                         int iy, int iz) {
  double v = outer_u(g, e, ix, iy, iz);
  if (cf != 0.0) v = v + cf * source(P, g, e, g.ind[node_index(g, ix, iy, iz)] > 0, ix, iy, iz);
  return v;
}

MTS_HD void op_weights(const PhysConst &P, const DevGeom &g, int d,
                                           bool outside, int ix, int iy, int iz, double vd[3]) {
  op_weights_with(P, d, outside, TabAt{g, {ix, iy, iz}}, vd);
}

// [u] and u_dir = P^{-1}[d][0] n.[beta grad u] at one crossing
// (Refresh_Jump, intersection.cc:2294-2490, approximate-jump mode).
MTS_HD void jumps(const double *cr, double e, double &ju, double &jb) {
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
MTS_HD double jl(const double *w, double juL, double jbL, double juR,
                                     double jbR) {
  return juL * w[5] + jbL * w[6] + juR * w[7] + jbR * w[8];
}

// ------------------------------------------- MIB fictitious-point stencils
// Sums of the patched rows without jump terms, in the reference's order
// (douglas_adi.cc:116-195).  F(i) returns the field value at line row i.
template <class F>
MTS_HD double irr_sum0(const double *v, F f, int ll) {
  const double *vd = v + kIrrVd, *wr = v + kIrrWr;
  double s = vd[0] * f(ll - 1) + vd[1] * f(ll);
  for (int i = 0; i < 4; i++) s += vd[2] * f(ll - 1 + i) * wr[i];
  return s;
}
template <class F>
MTS_HD double irr_sum1(const double *v, F f, int ll) {
  const double *vd = v + kIrrVd + 3, *wl = v + kIrrWl;
  double s = vd[1] * f(ll + 1) + vd[2] * f(ll + 2);
  for (int i = 0; i < 4; i++) s += vd[0] * f(ll - 1 + i) * wl[i];
  return s;
}
template <class F>
MTS_HD double cor_sum0(const double *v, F f, int ll) {
  const double *vd = v + kCorVd, *w = v + kCorLwr;
  double s = vd[0] * f(ll - 1) + vd[1] * f(ll);
  for (int i = 0; i < 5; i++) s += vd[2] * f(ll - 1 + i) * w[i];
  return s;
}
template <class F>
MTS_HD double cor_sum2(const double *v, F f, int ll) {
  const double *vd = v + kCorVd + 6, *w = v + kCorRwl;
  double s = vd[1] * f(ll + 2) + vd[2] * f(ll + 3);
  for (int i = 0; i < 5; i++) s += vd[0] * f(ll - 1 + i) * w[i];
  return s;
}
// middle corner row; with jumps interleaved as the Crank-Nicolson operators do
// (lod.cc:414-440) when cn is set, else the Douglas order.
template <class F>
MTS_HD double cor_sum1(const double *v, F f, int ll, bool cn, double jL,
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

// segment of lane for a line of n rows: p lanes active, rows balanced
MTS_HD void segment(int n, int lane, int &s, int &m) {
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

// src += jump terms of block row rr (Src_2nd, douglas_adi.cc:1123-1428)
MTS_HD double src_jump(const DevMib &mib, int bidx, int rr, double e,
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
template <typename FT>
MTS_HD double block_sum(const DevMib &mib, int bidx, int rr, const FT *u, size_t base,
                        size_t rs) {
  const BlockHdr h = mib.hdr[bidx];
  const double *v = mib.val + h.off;
  auto F = [&](int i) { return (double)u[base + (size_t)i * rs]; };
  if (h.type == kBlkIrr) return rr == 0 ? irr_sum0(v, F, h.ll) : irr_sum1(v, F, h.ll);
  if (rr == 0) return cor_sum0(v, F, h.ll);
  if (rr == 1) return cor_sum1(v, F, h.ll, false, 0.0, 0.0);
  return cor_sum2(v, F, h.ll);
}

// Douglas stage-1 right-hand side at one node (douglas_adi.cc:93-359):
// R = dt L_y u + dt L_z u + u + dt f, boundary nodes get the boundary value.
// Every load is issued before the first is used (a boundary node reads itself
// in place of its neighbours), so a node costs one memory round trip.
struct Rhs1Trig {
  double s[3], c[3], cym, cyp;
  MTS_HD double operator()(int a, int kind) const {
    return kind == kTabSin ? s[a] : kind == kTabCos ? c[a] : kind == kTabCosM ? cym : cyp;
  }
};
// R at an interior node from its value, its y and z neighbours and the
// node's trigonometric values (douglas_adi.cc:93-359 operation order).
MTS_HD double rhs1_value(const PhysConst &P, double dt, double e, bool out, const Rhs1Trig &t,
                         double u, double uym, double uyp, double uzm, double uzp) {
  double vy[3], vz[3];
  op_weights_with(P, 1, out, t, vy);
  op_weights_with(P, 2, out, t, vz);
  double sum_y = uym * vy[0] + u * vy[1] + uyp * vy[2];
  double sum_z = uzm * vz[0] + u * vz[1] + uzp * vz[2];
  double r = dt * sum_y;
  r += dt * sum_z;
  r += u;
  r += dt * source_with(P, e, out, t);
  return r;
}
template <typename F = double>
MTS_HD void rhs1_node(const Rhs1Args &p, int ix, int iy, int iz) {
  const DevGeom &g = p.g;
  const F *pu = static_cast<const F *>(p.u);
  F *pR = static_cast<F *>(p.R);
  const size_t c = node_index(g, ix, iy, iz);
  const bool bnd = ix == 0 || iy == 0 || iz == 0 || ix == g.n[0] - 1 || iy == g.n[1] - 1 ||
                   iz == g.n[2] - 1;
  const size_t sy = bnd ? 0 : (size_t)g.px, sz = bnd ? 0 : (size_t)g.n[1] * g.px;
  const signed char ind = g.ind[c];
  const double u = pu[c];
  const double uym = pu[c - sy], uyp = pu[c + sy], uzm = pu[c - sz], uzp = pu[c + sz];
  Rhs1Trig t;
  t.s[0] = tabv(g, 0, kTabSin, ix);
  t.s[1] = tabv(g, 1, kTabSin, iy);
  t.s[2] = tabv(g, 2, kTabSin, iz);
  t.c[0] = tabv(g, 0, kTabCos, ix);
  t.c[1] = tabv(g, 1, kTabCos, iy);
  t.c[2] = tabv(g, 2, kTabCos, iz);
  t.cym = tabv(g, 1, kTabCosM, iy);
  t.cyp = tabv(g, 1, kTabCosP, iy);
  if (bnd) {
    pR[c] = (F)outer_with(p.e, t);
    return;
  }
  pR[c] = (F)rhs1_value(p.P, p.dt, p.e, ind > 0, t, u, uym, uyp, uzm, uzp);
}

// The same for the node pair (2 pp, 2 pp + 1) of x row `row` (= iz * ny +
// iy), pp < px / 2: one 16-byte load per field row, one 16-byte store when
// both nodes are written.  Nodes flagged as MIB-affected (|ind| == 2, spec_node
// writes them) and the x padding are left alone.
template <typename F = double>
MTS_HD void rhs1_pair(const Rhs1Args &p, unsigned row, unsigned pp) {
  const DevGeom &g = p.g;
  const F *pu = static_cast<const F *>(p.u);
  F *pR = static_cast<F *>(p.R);
  const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  const int iy = (int)(row % (unsigned)n1), iz = (int)(row / (unsigned)n1);
  const int ix0 = 2 * (int)pp;
  if (ix0 >= n0) return;
  const bool two = ix0 + 1 < n0;
  const bool bnd_row = iy == 0 || iz == 0 || iy == n1 - 1 || iz == n2 - 1;
  const size_t c = node_index(g, ix0, iy, iz);
  const size_t sy = bnd_row ? 0 : (size_t)g.px, sz = bnd_row ? 0 : (size_t)n1 * g.px;
  // (the pitch is even and a whole number of 16-byte units, so c is 16-byte
  // aligned; the second element of a pair past the row's end is padding)
  double u[2], uym[2], uyp[2], uzm[2], uzp[2];
  ld2(pu + c, u);
  ld2(pu + c - sy, uym);
  ld2(pu + c + sy, uyp);
  ld2(pu + c - sz, uzm);
  ld2(pu + c + sz, uzp);
  const signed char i0 = g.ind[c], i1 = g.ind[c + 1];
  Rhs1Trig t;
  t.s[1] = tabv(g, 1, kTabSin, iy);
  t.s[2] = tabv(g, 2, kTabSin, iz);
  t.c[1] = tabv(g, 1, kTabCos, iy);
  t.c[2] = tabv(g, 2, kTabCos, iz);
  t.cym = tabv(g, 1, kTabCosM, iy);
  t.cyp = tabv(g, 1, kTabCosP, iy);
  const double sx[2] = {tabv(g, 0, kTabSin, ix0), two ? tabv(g, 0, kTabSin, ix0 + 1) : 0.0};
  const double cx[2] = {tabv(g, 0, kTabCos, ix0), two ? tabv(g, 0, kTabCos, ix0 + 1) : 0.0};
  double r[2];
  bool w[2];
  for (int q = 0; q < 2; q++) {
    const int ix = ix0 + q;
    const signed char ind = q == 0 ? i0 : i1;
    w[q] = ix < n0 && ind != 2 && ind != -2;
    t.s[0] = sx[q];
    t.c[0] = cx[q];
    const bool bnd = bnd_row || ix == 0 || ix == n0 - 1;
    r[q] = bnd ? outer_with(p.e, t)
               : rhs1_value(p.P, p.dt, p.e, ind > 0, t, u[q], uym[q], uyp[q], uzm[q], uzp[q]);
  }
  if (w[0] && w[1]) {
    pR[c] = (F)r[0];
    pR[c + 1] = (F)r[1];
  } else {
    if (w[0]) pR[c] = (F)r[0];
    if (w[1]) pR[c + 1] = (F)r[1];
  }
}

// Douglas stage-1 right-hand side at the t-th MIB-affected node: the
// fictitious-point stencils along y and z and the source jump terms of all
// three directions (douglas_adi.cc:116-312, Src_2nd :1158-1428).
template <typename F = double>
MTS_HD void spec_node(const SpecArgs &p, int t) {
  const DevGeom &g = p.g;
  const F *pu = static_cast<const F *>(p.u);
  const int *rec = p.spec + (size_t)t * kSpecLen;
  const int ix = rec[0], iy = rec[1], iz = rec[2];
  const size_t c = node_index(g, ix, iy, iz);
  const bool out = g.ind[c] > 0;
  const size_t sy = (size_t)g.px, sz = (size_t)g.n[1] * g.px;
  const double u = pu[c];
  double sum_y, sum_z;
  if (rec[3 + 1] >= 0) {
    sum_y = block_sum(p.mib[1], rec[3 + 1], rec[6 + 1], pu, c - (size_t)iy * sy, sy);
  } else {
    double vy[3];
    op_weights(p.P, g, 1, out, ix, iy, iz, vy);
    sum_y = (double)pu[c - sy] * vy[0] + u * vy[1] + (double)pu[c + sy] * vy[2];
  }
  if (rec[3 + 2] >= 0) {
    sum_z = block_sum(p.mib[2], rec[3 + 2], rec[6 + 2], pu, c - (size_t)iz * sz, sz);
  } else {
    double vz[3];
    op_weights(p.P, g, 2, out, ix, iy, iz, vz);
    sum_z = (double)pu[c - sz] * vz[0] + u * vz[1] + (double)pu[c + sz] * vz[2];
  }
  double src = source(p.P, g, p.e, out, ix, iy, iz);
  for (int d = 0; d < 3; d++)
    if (rec[3 + d] >= 0) src = src_jump(p.mib[d], rec[3 + d], rec[6 + d], p.e, src);
  double r = p.dt * sum_y;
  r += p.dt * sum_z;
  r += u;
  r += p.dt * src;
  static_cast<F *>(p.R)[c] = (F)r;
}

// ------------------------------------------------------- line stages
// A workgroup of 256 threads solves one tile of lines at once: 256 / L
// lines (16 lines of a 16-lane split, 8 of a 32-lane one, 4 of a 64-lane one).
constexpr int kBlockThreads = 256;
MTS_HD constexpr int tile_lines(int L) { return kBlockThreads / L; }

// Split of an n-node line over a lane group (see LinePart).  Picks the
// smallest group (16, then 32, then 64 lanes, at least min_lanes) and
// register bound (4, 8, 16 rows per lane) that hold the interior rows with no
// segment of one row.
inline bool make_line_part(int n, LinePart &q, int min_lanes = 16) {
  const int ni = n - 2;
  if (ni < 2 || n > 1024) return false;
  static const int kL[3] = {16, 32, 64};
  static const int kM[3] = {4, 8, 16};
  for (int li = 0; li < 3; li++) {
    const int L = kL[li];
    if (L < min_lanes) continue;
    int m = (ni + L - 1) / L;
    if (m < 2) m = 2;
    int p = (ni + m - 1) / m;
    int mlast = ni - (p - 1) * m;
    if (mlast == 1) {  // fold a one-row tail into the previous segment
      p -= 1;
      mlast = m + 1;
    }
    if (p > L) continue;
    const int mx = m > mlast ? m : mlast;
    for (int mi = 0; mi < 3; mi++) {
      if (mx > kM[mi]) continue;
      q.n = n;
      q.L = L;
      q.mmax = kM[mi];
      q.m = m;
      q.mlast = mlast;
      q.p = p;
      q.stride = mx | 1;
      q.magic = ((1 << 20) + m - 1) / m;
      q.lds_len = (2 + p * q.stride) | 1;  // odd: neighbouring lines on other banks
      return true;
    }
  }
  return false;
}

// LDS slot of row r of a line
// (selects rather than early returns: the kernels form it per element)
MTS_HD int lds_slot(const LinePart &q, int r) {
  const int t = r - 1;
  int k = (int)(((unsigned)t * (unsigned)q.magic) >> 20);  // t / m for 0 <= t < 2^20 / m
  k = k < q.p - 1 ? k : q.p - 1;
  const int mid = 1 + k * q.stride + (t - k * q.m);
  return r <= 0 ? 0 : (r >= q.n - 1 ? 1 + q.p * q.stride : mid);
}

// Segment of lane kseg of a group: first row s, m rows (0: idle lane).
struct LaneSeg {
  int k, s, m;
};
MTS_HD LaneSeg lane_seg(const LinePart &q, int kseg, bool line_valid) {
  LaneSeg r;
  r.k = kseg;
  if (!line_valid || kseg >= q.p) {
    r.s = q.n - 1;
    r.m = 0;
  } else {
    r.s = 1 + kseg * q.m;
    r.m = kseg == q.p - 1 ? q.mlast : q.m;
  }
  return r;
}

template <int DIR>
struct LineMap {
  // number of lines along the tile index j and the outer index k
  MTS_HD static int nj(const DevGeom &g) { return DIR == 0 ? g.n[1] : g.n[0]; }
  MTS_HD static int nk(const DevGeom &g) { return DIR == 2 ? g.n[1] : g.n[2]; }
  MTS_HD static size_t stride(const DevGeom &g) {
    return DIR == 0 ? 1 : DIR == 1 ? (size_t)g.px : (size_t)g.n[1] * g.px;
  }
  MTS_HD static size_t base(const DevGeom &g, int j, int k) {
    return DIR == 0 ? ((size_t)k * g.n[1] + j) * g.px
                    : DIR == 1 ? (size_t)k * g.n[1] * g.px + j : (size_t)k * g.px + j;
  }
  MTS_HD static void node(int j, int k, int i, int &ix, int &iy, int &iz) {
    if (DIR == 0) { ix = i; iy = j; iz = k; }
    else if (DIR == 1) { ix = j; iy = i; iz = k; }
    else { ix = j; iy = k; iz = i; }
  }
};

// The source f at the rows of one line.  Along a z line x and y are fixed,
// so f is sin z (inside) or cos z (outside) times a factor of the line:
// f = e sz [(3 beta - 1) sx sy - beta_x cx sy - beta_y sx cy] inside and
// f = e (3 beta+ - 1) cx cy cz outside (the factors of source(), regrouped).
// Other directions evaluate source() per row.
// Along z lines sin z and cos z come from zt (sin z at rows 0..n-1, then
// cos z), a copy of the axis tables staged in LDS, or the tables themselves
// when zt is null.
template <int DIR>
struct LineSource {
  double fin = 0.0, fout = 0.0;
  const double *zs = nullptr, *zc = nullptr;
  MTS_HD LineSource() {}
  MTS_HD LineSource(const PhysConst &P, const DevGeom &g, int j, int k, const double *zt) {
    if (DIR == 2) {
      zs = zt ? zt : g.tab[2] + kTabSin * g.n[2];
      zc = zt ? zt + g.n[2] : g.tab[2] + kTabCos * g.n[2];
      const double sx = tabv(g, 0, kTabSin, j), sy = tabv(g, 1, kTabSin, k);
      const double cx = tabv(g, 0, kTabCos, j), cy = tabv(g, 1, kTabCos, k);
      fout = P.f_out * cx * cy;
      if (P.beta_kind == 0) {
        fin = P.f_in * sx * sy;
      } else {
        const double b = P.b_in * (1.0 + 0.5 * sx * cy);
        const double bx = 0.5 * P.b_in * cx * cy;
        const double by = -0.5 * P.b_in * sx * sy;
        fin = (3.0 * b - 1.0) * sx * sy - bx * cx * sy - by * sx * cy;
      }
    }
  }
  MTS_HD double at(const PhysConst &P, const DevGeom &g, double e, bool out, int j, int k,
                   int i) const {
    if (DIR == 2) return e * (out ? fout * zc[i] : fin * zs[i]);
    int ix, iy, iz;
    LineMap<DIR>::node(j, k, i, ix, iy, iz);
    return source(P, g, e, out, ix, iy, iz);
  }
};

// End row e of the tile's lines (line e / 2; row 0 for even e, n - 1 for
// odd): the analytic boundary value at the stage's input time, for stages
// whose input end rows are boundary nodes the previous stage did not write.
// The tile walk skips these rows when in_analytic is set (their loaded values
// would be replaced), so the kernel computes the values before the tile
// loads and stores them with the tile, with no ordering between the two.
// tile_end_slot: the LDS slot of end row e, or -1 if its line lies past the grid.
template <int DIR>
MTS_HD int tile_end_slot(const LineArgs &p, int j0, int e) {
  const int n = p.g.n[DIR];
  const int l = e >> 1;
  if (j0 + l >= LineMap<DIR>::nj(p.g)) return -1;
  return l * p.q.lds_len + lds_slot(p.q, (e & 1) ? n - 1 : 0);
}
template <int DIR>
MTS_HD double tile_end_val(const LineArgs &p, int j0, int k, int e) {
  const int n = p.g.n[DIR];
  const int j = j0 + (e >> 1);
  if (j >= LineMap<DIR>::nj(p.g)) return 0.0;
  int ix, iy, iz;
  LineMap<DIR>::node(j, k, (e & 1) ? n - 1 : 0, ix, iy, iz);
  return outer_u(p.g, p.e_in, ix, iy, iz);
}
template <int DIR, typename R>
MTS_HD void tile_end_value(const LineArgs &p, int j0, int k, int e, R *tA) {
  const int sl = tile_end_slot<DIR>(p, j0, e);
  if (sl >= 0) tA[sl] = (R)tile_end_val<DIR>(p, j0, k, e);
}
// The Douglas stage-1 right-hand side computed inside the x sweep
// (kFlagPre1): the sweep loads u^n and turns every interior element of its
// tile into R as it stages it, so R is never written to or read from memory.
// Along an x line (j = iy, k = iz) the y and z factors are fixed.
struct Pre1Line {
  double sy, cy, cym, cyp, sz, cz;
};
MTS_HD Pre1Line pre1_line(const DevGeom &g, int j, int k) {
  Pre1Line l;
  l.sy = tabv(g, 1, kTabSin, j);
  l.cy = tabv(g, 1, kTabCos, j);
  l.cym = tabv(g, 1, kTabCosM, j);
  l.cyp = tabv(g, 1, kTabCosP, j);
  l.sz = tabv(g, 2, kTabSin, k);
  l.cz = tabv(g, 2, kTabCos, k);
  return l;
}
// Elements per thread that cover a tile: T * n <= 17 * 256 for every split
// make_line_part returns.
constexpr int kTileElsPerThread = 17;

// Pairs per thread that cover a tile: T * ceil(n / 2) <= 9 * 256.
constexpr int kTilePairsPerThread = 9;
// The kernels move them kTileChunk at a time, so steps u = 0 .. kTileSteps - 1
// run (the last ones find no element).
constexpr int kTileChunk = 9;
constexpr int kTileSteps = (kTilePairsPerThread + kTileChunk - 1) / kTileChunk * kTileChunk;

// The element pairs one thread moves between a tile and HBM, u = 0, 1, ...
// (kTilePairsPerThread steps cover the tile), with everything that depends on
// the thread only formed once per tile.  Both elements of a pair are
// neighbours in memory (the second at global index gi + 1), so a pair is one
// 16-byte access whenever the first exists.  The second may not (a line past
// the grid, the row past an odd line's end); the 16 bytes are then still
// inside the padded x row (the pitch is a whole number of 128-byte lines,
// above n when n is odd), so the access stays inside the field.
//   y/z lines: T / 2 pairs per tile row; the thread always moves the same two
//   neighbouring lines (2 (tid mod T/2) and the next) and its row advances by
//   kBlockThreads / (T / 2) per step.
//   x lines: L = kBlockThreads / T threads per line; thread tid moves rows
//   2 (tid mod L) + 2 L u and the next of line tid / L.
struct WalkPair {
  size_t gi;       // global index of the first element
  int sl0, sl1;    // LDS slots of the two elements
  bool ok0, ok1;   // the element exists (line inside the grid, row inside the line)
  bool w0, w1;     // and is written back (an interior line)
  bool ld0, ld1;   // and is loaded (not an input end row of an in_analytic stage)
  int r;           // row of the first element
};
template <int DIR, int T>
struct TileWalk {
  static constexpr int kL = kBlockThreads / T;  // x lines: threads per line
  static constexpr int kRowStep = DIR == 0 ? 2 * kL : kBlockThreads / (T / 2);
  size_t gb;          // global index of row 0 of the thread's (first) line
  int sl;             // LDS base of that line
  int r0;             // the thread's row at u = 0
  int j;              // the thread's (first) line
  bool okl0, okl1;    // its lines lie inside the grid (x: one line, twice)
  bool wl0, wl1;      // and are interior lines
  MTS_HD TileWalk(const LineArgs &p, int j0, int k, int tid) {
    using M = LineMap<DIR>;
    const int nj = M::nj(p.g);
    int l;
    if (DIR == 0) {
      l = tid / kL;
      r0 = 2 * (tid % kL);
      j = j0 + l;
      okl0 = okl1 = j < nj;
      wl0 = wl1 = j >= 1 && j <= nj - 2;
    } else {
      l = 2 * (tid & (T / 2 - 1));
      r0 = tid / (T / 2);
      j = j0 + l;
      okl0 = j < nj;
      okl1 = j + 1 < nj;
      wl0 = j >= 1 && j <= nj - 2;
      wl1 = j + 1 <= nj - 2;
    }
    gb = M::base(p.g, j, k);
    sl = l * p.q.lds_len;
  }
  MTS_HD WalkPair at(const LineArgs &p, int u) const {
    using M = LineMap<DIR>;
    const int n = p.g.n[DIR];
    WalkPair w;
    w.r = r0 + u * kRowStep;
    if (DIR == 0) {
      w.ok0 = okl0 && w.r < n;
      w.ok1 = okl0 && w.r + 1 < n;
      w.gi = gb + (size_t)w.r;
      w.sl0 = sl + lds_slot(p.q, w.r);
      w.sl1 = sl + lds_slot(p.q, w.r + 1);
    } else {
      const bool in = w.r < n;
      w.ok0 = okl0 && in;
      w.ok1 = okl1 && in;
      w.gi = gb + (size_t)w.r * M::stride(p.g);
      w.sl0 = sl + lds_slot(p.q, w.r);
      w.sl1 = w.sl0 + p.q.lds_len;
    }
    // the write-back covers the end rows either way (the stage's output)
    w.w0 = w.ok0 && wl0;
    w.w1 = w.ok1 && wl1;
    if (p.in_analytic) {
      // input end rows come from tile_end_val instead
      const int r1 = DIR == 0 ? w.r + 1 : w.r;
      w.ld0 = w.ok0 && w.r != 0 && w.r != n - 1;
      w.ld1 = w.ok1 && r1 != 0 && r1 != n - 1;
    } else {
      w.ld0 = w.ok0;
      w.ld1 = w.ok1;
    }
    return w;
  }
};

// The loads of one element pair (rows r, r + 1 of the line, global index gi,
// 16-byte aligned): u^n, its y and z neighbours, the classification and the
// x tables.  Only for pairs of interior lines (the neighbours exist).
struct Pre1Pair {
  double u[2], ym[2], yp[2], zm[2], zp[2], sx[2], cx[2];
  signed char ind[2];
};
MTS_HD void pre1_load(const LineArgs &p, size_t gi, int r, Pre1Pair &e) {
  const DevGeom &g = p.g;
  const double *u = static_cast<const double *>(p.A);
  const size_t sy = (size_t)g.px, sz = (size_t)g.n[1] * g.px;
  ld2(u + gi, e.u);
  ld2(u + gi - sy, e.ym);
  ld2(u + gi + sy, e.yp);
  ld2(u + gi - sz, e.zm);
  ld2(u + gi + sz, e.zp);
  e.ind[0] = g.ind[gi];
  e.ind[1] = g.ind[gi + 1];
  const int n = g.n[0];
  e.sx[0] = tabv(g, 0, kTabSin, r);
  e.cx[0] = tabv(g, 0, kTabCos, r);
  e.sx[1] = r + 1 < n ? tabv(g, 0, kTabSin, r + 1) : 0.0;
  e.cx[1] = r + 1 < n ? tabv(g, 0, kTabCos, r + 1) : 0.0;
}
// R of element q of the pair: rhs1_value in the stage-1 launch's operation
// order (the same value, bit for bit), or at a MIB-affected node
// (classification +-2) the value spec_node wrote into p.S.
MTS_HD double pre1_value(const LineArgs &p, const Pre1Line &l, const Pre1Pair &e, size_t gi,
                         int q) {
  if (e.ind[q] == 2 || e.ind[q] == -2) return static_cast<const double *>(p.S)[gi + q];
  Rhs1Trig t;
  t.s[0] = e.sx[q];
  t.c[0] = e.cx[q];
  t.s[1] = l.sy;
  t.c[1] = l.cy;
  t.s[2] = l.sz;
  t.c[2] = l.cz;
  t.cym = l.cym;
  t.cyp = l.cyp;
  return rhs1_value(p.P, p.dt, p.e_in, e.ind[q] > 0, t, e.u[q], e.ym[q], e.yp[q], e.zm[q],
                    e.zp[q]);
}

// ------------------------------------------------ interface diagnostics
// The exact jumps at crossing c (douglas_adi.cc's equation: u- = e sin sin
// sin, u+ = e cos cos cos): [u], the GFM's u_dir (P^{-1}[d][0] times row 0 of
// P times [beta grad u]) and the exact [beta u_d], in the operation order of
// the host diagnostics and of jumps().
template <int DIR>
MTS_HD void diag_jumps(const double *cr, double e, double &ju, double &udir, double &bd) {
  const double sx = cr[0], sy = cr[1], sz = cr[2], cx = cr[3], cy = cr[4], cz = cr[5];
  const double bi = cr[6], bo = cr[7];
  const double me = -e;
  ju = e * cx * cy * cz - e * sx * sy * sz;
  const double jbx = bo * (me * sx * cy * cz) - bi * (e * cx * sy * sz);
  const double jby = bo * (me * cx * sy * cz) - bi * (e * sx * cy * sz);
  const double jbz = bo * (me * cx * cy * sz) - bi * (e * sx * sy * cz);
  const double xi = cr[8] * jbx + cr[9] * jby + cr[10] * jbz;
  udir = cr[11] * xi;
  bd = DIR == 0 ? jbx : (DIR == 1 ? jby : jbz);
}
// The errors of every crossing of line lid (Refresh_Fp + Error_Fp,
// intersection.cc:1841-2281 and 4667-4771; Refresh_Jump's jump.err and
// Error_Jump, :2294-2490 and 4594-4665): each fictitious point from its
// weights (the block record's copies of the crossing's), the exact solution
// on the line's nodes and the exact jumps, relative to the other side's
// exact value; the jump error |u_dir - [beta u_d]| / [beta u_d].  A corner
// pair shares its middle fictitious point (errr of the first crossing is errl
// of the second).
template <int DIR>
MTS_HD void diag_line(const DiagArgs &a, int lid) {
  const int bb = a.mib.blk_ptr[lid], be = a.mib.blk_ptr[lid + 1];
  const int j = lid % a.nj, k = lid / a.nj;
  const DevGeom &g = a.g;
  const double e = a.e;
  // the exact value at line node i on its own side (inv: the other side)
  auto node_u = [&](int i, bool inv) {
    int ix, iy, iz;
    LineMap<DIR>::node(j, k, i, ix, iy, iz);
    const bool out = g.ind[node_index(g, ix, iy, iz)] > 0;
    const double outer = e * tabv(g, 0, kTabCos, ix) * tabv(g, 1, kTabCos, iy) *
                         tabv(g, 2, kTabCos, iz);
    const double inner = e * tabv(g, 0, kTabSin, ix) * tabv(g, 1, kTabSin, iy) *
                         tabv(g, 2, kTabSin, iz);
    return out != inv ? outer : inner;
  };
  auto rel = [](double ref, double v) { return fabs((ref - v) / ref); };
  for (int bi = bb; bi < be; bi++) {
    const BlockHdr h = a.mib.hdr[bi];
    const double *v = a.mib.val + h.off;
    const int ll = h.ll, c = h.c0;
    double juL, udL, bdL;
    diag_jumps<DIR>(a.cr + (size_t)c * kCrLen, e, juL, udL, bdL);
    a.je[c] = fabs(udL - bdL) / bdL;
    if (h.type == kBlkIrr) {
      const double *wl = v + kIrrWl, *wr = v + kIrrWr;
      double sl = 0.0, sr = 0.0;
      sl += wl[4] * juL;
      sl += wl[5] * bdL;
      sr += wr[4] * juL;
      sr += wr[5] * bdL;
      for (int q = 0; q < 4; q++) {
        const double u = node_u(ll - 1 + q, false);
        sl += wl[q] * u;
        sr += wr[q] * u;
      }
      a.el[c] = rel(node_u(ll, true), sl);
      a.er[c] = rel(node_u(ll + 1, true), sr);
    } else {
      double juR, udR, bdR;
      diag_jumps<DIR>(a.cr + (size_t)(c + 1) * kCrLen, e, juR, udR, bdR);
      a.je[c + 1] = fabs(udR - bdR) / bdR;
      const double *lwl = v + kCorLwl, *lwr = v + kCorLwr, *rwr = v + kCorRwr;
      double lsl = 0.0, lsr = 0.0, rsr = 0.0;
      lsl += lwl[5] * juL;
      lsl += lwl[6] * bdL;
      lsl += lwl[7] * juR;
      lsl += lwl[8] * bdR;
      lsr += lwr[5] * juL;
      lsr += lwr[6] * bdL;
      lsr += lwr[7] * juR;
      lsr += lwr[8] * bdR;
      rsr += rwr[5] * juL;
      rsr += rwr[6] * bdL;
      rsr += rwr[7] * juR;
      rsr += rwr[8] * bdR;
      for (int q = 0; q < 5; q++) {
        const double u = node_u(ll - 1 + q, false);
        lsl += lwl[q] * u;
        lsr += lwr[q] * u;
        rsr += rwr[q] * u;
      }
      const double mid = rel(node_u(ll + 1, true), lsr);
      a.el[c] = rel(node_u(ll, true), lsl);
      a.er[c] = mid;
      a.el[c + 1] = mid;
      a.er[c + 1] = rel(node_u(ll + 2, true), rsr);
    }
  }
}

// Per-line table entries a lane needs, read once up front.
struct LineMeta {
  int bb, be;    // block range
  bool out0;     // first node outside
  bool seq;      // solved with the sequential Thomas algorithm
};
MTS_HD LineMeta line_meta(const DevMib &mib, int lid, bool valid) {
  LineMeta m;
  m.bb = m.be = 0;
  m.out0 = false;
  m.seq = false;
  if (valid) {
    m.bb = mib.blk_ptr[lid];
    m.be = mib.blk_ptr[lid + 1];
    m.out0 = mib.cls0[lid] != 0;
    m.seq = mib.seq != nullptr && mib.seq[lid] != 0;
  }
  return m;
}

// Classification (true = outside) of row r of line lid: the line's first
// node class flipped at every crossing left of r (an irregular block holds
// one crossing at ll, a corner pair two, at ll and ll + 1).
MTS_HD bool row_outside(const DevMib &mib, const LineMeta &lm, int r) {
  bool o = lm.out0;
  for (int bi = lm.bb; bi < lm.be; bi++) {
    const BlockHdr h = mib.hdr[bi];
    if (h.ll < r) o = !o;
    if (h.type == kBlkCor && h.ll + 1 < r) o = !o;
  }
  return o;
}
// Classification bits of rows s .. s + 31 (bit t set: row s + t outside).
MTS_HD unsigned rows_outside(const DevMib &mib, const LineMeta &lm, int s) {
  unsigned mask = lm.out0 ? 0xffffffffu : 0u;
  for (int bi = lm.bb; bi < lm.be; bi++) {
    const BlockHdr h = mib.hdr[bi];
    const int nf = h.type == kBlkCor ? 2 : 1;
    for (int f = 0; f < nf; f++) {
      const int t0 = h.ll + f - s + 1;  // rows s + t with t >= t0 lie right of the crossing
      if (t0 <= 0) mask = ~mask;
      else if (t0 < 32) mask ^= ~((1u << t0) - 1u);
    }
  }
  return mask;
}

// Right-hand side of the patched rows of one MIB block (douglas_adi.cc:116-195
// and :405-482, lod.cc:414-440 and :887-1078): fictitious-point stencil sums,
// jump terms and the Convert2Tri row reduction.  The block's rows ll-1 ..
// ll+3 of the input field(s) are read from LDS once (row indices past the
// line's end read its last row, unused).
struct BlockRhs {
  double v0, v1, v2;
};
template <int DIR, int FLAGS>
MTS_HD BlockRhs block_rhs(const LineArgs &p, int j, int k, const LineMeta &lm,
                          const BlockHdr h, const double *la, const double *lb) {
  constexpr int RHS = FLAGS & 3;
  constexpr bool LJ = (FLAGS & kFlagLhsJump) != 0;
  constexpr bool SOLVE = (FLAGS & kFlagSolve) != 0;
  using M = LineMap<DIR>;
  const DevGeom &g = p.g;
  const LinePart &q = p.q;
  const double tdt = p.theta_dt;
  const int ll = h.ll;
  double fa[5], fs[5];
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const int sl = lds_slot(q, ll - 1 + t);
    fa[t] = la[sl];
    fs[t] = RHS == kRhsDouglas ? lb[sl] : 0.0;
  }
  auto FA = [&](int i) { return fa[i - (ll - 1)]; };
  auto FS = [&](int i) { return fs[i - (ll - 1)]; };
  const double *v = p.mib.val + h.off;
  double ju = 0.0, jb = 0.0, juR = 0.0, jbR = 0.0;
  if (RHS == kRhsCN || RHS == kRhsTSZ || LJ) {
    jumps(p.mib.cr + (size_t)h.c0 * kCrLen, p.e_jump, ju, jb);
    if (h.type == kBlkCor) jumps(p.mib.cr + (size_t)(h.c0 + 1) * kCrLen, p.e_jump, juR, jbR);
  }
  double bv[3] = {0.0, 0.0, 0.0};
  if (h.type == kBlkIrr) {
    const double *vd0 = v + kIrrVd, *vd1 = v + kIrrVd + 3;
    const double *wl = v + kIrrWl, *wr = v + kIrrWr;
    const double J0 = ju * wr[4] + jb * wr[5];
    const double J1 = ju * wl[4] + jb * wl[5];
    if (RHS == kRhsCopy) {
      bv[0] = FA(ll);
      bv[1] = FA(ll + 1);
    } else if (RHS == kRhsDouglas) {
      bv[0] = -p.dt * irr_sum0(v, FS, ll) + FA(ll);
      bv[1] = -p.dt * irr_sum1(v, FS, ll) + FA(ll + 1);
    } else {
      double s0 = irr_sum0(v, FA, ll) + vd0[2] * J0;
      MTS_PHASE();
      double s1 = irr_sum1(v, FA, ll) + vd1[0] * J1;
      MTS_PHASE();
      bv[0] = tdt * s0 + FA(ll);
      bv[1] = tdt * s1 + FA(ll + 1);
    }
    if (RHS == kRhsTSZ) {
      for (int r = 0; r < 2; r++) {
        int ix, iy, iz;
        M::node(j, k, ll + r, ix, iy, iz);
        const bool out = row_outside(p.mib, lm, ll + r);
        bv[r] = bv[r] + p.c_f1 * source(p.P, g, p.e_f1, out, ix, iy, iz);
        bv[r] = bv[r] + p.c_f2 * source(p.P, g, p.e_f2, out, ix, iy, iz);
      }
    }
    if (LJ) {
      bv[0] = bv[0] + tdt * vd0[2] * J0;
      bv[1] = bv[1] + tdt * vd1[0] * J1;
    }
    if (SOLVE) {  // Convert2Tri row reduction, only where the rows are solved
      const double *mu = v + kIrrMul;
      bv[1] -= mu[0] * bv[0];
      bv[0] -= mu[1] * bv[1];
    }
  } else {
    const double *vd0 = v + kCorVd, *vd1 = v + kCorVd + 3, *vd2 = v + kCorVd + 6;
    const double JLwr = jl(v + kCorLwr, ju, jb, juR, jbR);
    const double JLwl = jl(v + kCorLwl, ju, jb, juR, jbR);
    const double JRwr = jl(v + kCorRwr, ju, jb, juR, jbR);
    const double JRwl = jl(v + kCorRwl, ju, jb, juR, jbR);
    MTS_PHASE();
    if (RHS == kRhsCopy) {
      bv[0] = FA(ll);
      bv[1] = FA(ll + 1);
      bv[2] = FA(ll + 2);
    } else if (RHS == kRhsDouglas) {
      bv[0] = -p.dt * cor_sum0(v, FS, ll) + FA(ll);
      MTS_PHASE();
      bv[1] = -p.dt * cor_sum1(v, FS, ll, false, 0.0, 0.0) + FA(ll + 1);
      MTS_PHASE();
      bv[2] = -p.dt * cor_sum2(v, FS, ll) + FA(ll + 2);
      MTS_PHASE();
    } else {
      double s0 = cor_sum0(v, FA, ll) + vd0[2] * JLwr;
      MTS_PHASE();
      double s1 = cor_sum1(v, FA, ll, true, JLwl, JRwr);
      MTS_PHASE();
      double s2 = cor_sum2(v, FA, ll) + vd2[0] * JRwl;
      MTS_PHASE();
      bv[0] = tdt * s0 + FA(ll);
      bv[1] = tdt * s1 + FA(ll + 1);
      bv[2] = tdt * s2 + FA(ll + 2);
    }
    if (RHS == kRhsTSZ) {
      for (int r = 0; r < 3; r++) {
        int ix, iy, iz;
        M::node(j, k, ll + r, ix, iy, iz);
        const bool out = row_outside(p.mib, lm, ll + r);
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
    if (SOLVE) {
      const double *mu = v + kCorMul;
      bv[1] -= mu[0] * bv[0];
      bv[2] -= mu[1] * bv[0];
      bv[0] -= mu[2] * bv[2];
      bv[1] -= mu[3] * bv[2];
      bv[2] -= mu[4] * bv[1];
      bv[0] -= mu[5] * bv[1];
    }
  }
  BlockRhs r;
  r.v0 = bv[0];
  r.v1 = bv[1];
  r.v2 = bv[2];
  return r;
}

// The block's five input rows ll-1 .. ll+3 from LDS (0 past the line's end).
template <int FLAGS, typename R>
MTS_HD void block_inputs(const LineArgs &p, int ll, const R *la, const R *lb,
                         double (&fa)[5], double (&fs)[5]) {
  constexpr bool HAS_S = (FLAGS & 3) == kRhsDouglas;
#pragma unroll
  for (int t = 0; t < 5; t++) {
    const int r = ll - 1 + t;
    const bool in = r <= p.q.n - 1;
    const int sl = lds_slot(p.q, r);
    fa[t] = in ? (double)la[sl] : 0.0;
    fs[t] = HAS_S && in ? (double)lb[sl] : 0.0;
  }
}
// Right-hand side of one patched row from its stage record.
template <int FLAGS>
MTS_HD double block_row_rhs(const LineArgs &p, const double *w, const double (&fa)[5],
                            const double (&fs)[5]) {
  constexpr int RHS = FLAGS & 3;
  double v = w[kRecWA] * fa[0];
#pragma unroll
  for (int t = 1; t < 5; t++) v += w[kRecWA + t] * fa[t];
  if (RHS == kRhsDouglas) {
#pragma unroll
    for (int t = 0; t < 5; t++) v += w[kRecWS + t] * fs[t];
  }
  v += p.e_jump * w[kRecK];
  if (RHS == kRhsTSZ) {
    v += p.e_f1 * w[kRecK1];
    v += p.e_f2 * w[kRecK2];
  }
  return v;
}

// Operator weights along the lines of one stage.  What is fixed along a line
// sits in a per-tile LDS table of kWtLen doubles per line, written once per
// tile (line_weight_entry): the outside weights (entries 0..2) and the inside
// ones (3..5: constant beta, or along z lines the variable beta's value, which
// depends on x and y only).  A row then reads its side's three weights from
// LDS at an address chosen by its side, rather than selecting between
// registers.  Inside a variable beta b_in (1 + 0.5 sin x cos y) along x and y
// lines, one factor is fixed along the line (0.5 cos y, 0.5 sin x: scaling by
// 0.5 is exact, so the products equal op_weights' bit for bit) and the other
// comes from the axis tables at the row.
constexpr int kWtLen = 6;
template <int DIR>
MTS_HD double line_weight_value(const LineArgs &p, int j0, int k, int e) {
  const int l = e / kWtLen, c = e % kWtLen;
  const int j = j0 + l;
  double v = 0.0;
  if (j < LineMap<DIR>::nj(p.g)) {
    if (c < 3) {
      v = p.P.vdc[DIR][1][c];
    } else if (DIR == 2 && p.P.beta_kind != 0) {
      double vd[3];
      op_weights(p.P, p.g, 2, false, j, k, 0, vd);
      v = vd[c - 3];
    } else {
      v = p.P.vdc[DIR][0][c - 3];
    }
  }
  return v;
}
template <int DIR>
MTS_HD void line_weight_entry(const LineArgs &p, int j0, int k, int e, double *wt) {
  wt[e] = line_weight_value<DIR>(p, j0, k, e);
}
// Entry e of the tile's copy of the z axis tables (LineSource): sin z at
// rows 0..n-1, then cos z.
MTS_HD double ztab_value(const DevGeom &g, int e) {
  const int n = g.n[2];
  return e < n ? tabv(g, 2, kTabSin, e) : tabv(g, 2, kTabCos, e - n);
}
MTS_HD void ztab_entry(const DevGeom &g, int e, double *zt) { zt[e] = ztab_value(g, e); }
template <int DIR>
struct LineWeights {
  const double *wl;    // the line's entries of the LDS table
  const double *tm, *tp;  // x, y lines, variable beta: axis tables at the lane's first row
  double hf;           // and the fixed half factor
  bool var;            // variable beta along the line (x, y lines)
  MTS_HD LineWeights(const PhysConst &P, const DevGeom &g, const double *wl_, int j, int s)
      : wl(wl_) {
    tm = tp = nullptr;
    hf = 0.0;
    var = DIR != 2 && P.beta_kind != 0;
    if (var) {
      const int a = DIR == 0 ? 0 : 1;
      hf = 0.5 * (DIR == 0 ? tabv(g, 1, kTabCos, j) : tabv(g, 0, kTabSin, j));
      tm = g.tab[a] + (DIR == 0 ? kTabSinM : kTabCosM) * g.n[a] + s;
      tp = g.tab[a] + (DIR == 0 ? kTabSinP : kTabCosP) * g.n[a] + s;
    }
  }
  // weights at row s + r of the line on the given side (the variable inside
  // weights are formed only where a lane needs them: a branch, which a
  // wavefront of outside rows skips)
  MTS_HD void at(const PhysConst &P, bool out, int r, double vd[3]) const {
    const double *w = wl + (out ? 0 : 3);
    vd[0] = w[0];
    vd[1] = w[1];
    vd[2] = w[2];
    if (DIR != 2 && var && !out) {
      const double b1 = P.b_in * (1.0 + tm[r] * hf), b2 = P.b_in * (1.0 + tp[r] * hf);
      vd[0] = b1 * P.ihh[DIR];
      vd[1] = -(b1 + b2) * P.ihh[DIR];
      vd[2] = b2 * P.ihh[DIR];
    }
  }
};

// Rows [s, s + m) of line (j, k): tridiagonal coefficients and right-hand
// side of the stage, including the line's MIB blocks (the reference's
// per-line assembly, e.g. douglas_adi.cc:377-491 and lod.cc:887-1078).  The
// couplings of rows 1 and n-2 to the boundary rows stay in a and c.
// la/lb: the line's input field(s) staged in LDS (slot layout of LinePart);
// wl: the line's entries of the tile's weight table (LineWeights); zt: see
// LineSource.
// R is the field type (double, or float for the FP32 form): the rows are
// formed in FP64 and rounded to R once.
template <int DIR, int MMAX, int FLAGS, typename R = double>
MTS_HD void line_rows(const LineArgs &p, int j, int k, const LineMeta &lm, const R *la,
                      const R *lb, const double *wl, const double *zt, const LaneSeg &sg,
                      R (&a)[MMAX], R (&b)[MMAX], R (&c)[MMAX], R (&d)[MMAX]) {
  constexpr int RHS = FLAGS & 3;
  constexpr bool SOLVE = (FLAGS & kFlagSolve) != 0;
  const DevGeom &g = p.g;
  const LinePart &q = p.q;
  const double tdt = p.theta_dt;
  const int s = sg.s, m = sg.m;
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    a[r] = 0.0;
    b[r] = 1.0;
    c[r] = 0.0;
    d[r] = 0.0;
  }
  if (m == 0) return;
  const unsigned cm = rows_outside(p.mib, lm, s);
  const int base = 1 + sg.k * q.stride;
  // the rows just before and after the segment
  const int sl_prev = sg.k == 0 ? 0 : base - q.stride + q.m - 1;
  const int sl_next = sg.k == q.p - 1 ? 1 + q.p * q.stride : base + q.stride;
  const LineSource<DIR> src(p.P, g, j, k, zt);
  (void)src;
  const LineWeights<DIR> lw(p.P, g, wl, j, s);
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    if (r < m) {
      const int i = s + r;
      const bool out = ((cm >> r) & 1u) != 0;
      double vd[3];
      lw.at(p.P, out, r, vd);
      // regular rows in the field's arithmetic (FP64, or FP32 for the FP32
      // form, whose rows are FP32 values either way)
      const R w0 = (R)vd[0], w1 = (R)vd[1], w2 = (R)vd[2];
      const R rt = (R)tdt;
      a[r] = -rt * w0;
      b[r] = R(1) - rt * w1;
      c[r] = -rt * w2;
      if (RHS == kRhsCopy) {
        d[r] = la[base + r];
      } else {
        const R *f = RHS == kRhsDouglas ? lb : la;
        const R fp = r == 0 ? f[sl_prev] : f[base + r - 1];
        const R fc = f[base + r];
        const R fn = r + 1 < m ? f[base + r + 1] : f[sl_next];
        // (the right-hand side with fused multiply-adds; the coefficients
        // above round as the reference's do)
        const R sum = fmad(fn, w2, fmad(fc, w1, fp * w0));
        if (RHS == kRhsDouglas) {
          d[r] = fmad(-(R)p.dt, sum, la[base + r]);
        } else {
          R v = fmad(rt, sum, fc);
          if (RHS == kRhsTSZ) {
            v = v + (R)(p.c_f1 * src.at(p.P, g, p.e_f1, out, j, k, i));
            v = v + (R)(p.c_f2 * src.at(p.P, g, p.e_f2, out, j, k, i));
          }
          d[r] = v;
        }
      }
    }
  }

  // ---- MIB blocks of this line
  for (int bi = lm.bb; bi < lm.be; bi++) {
    const BlockHdr h = p.mib.hdr[bi];
    const int ll = h.ll;
    const int nr = (h.type == kBlkIrr) ? 2 : 3;
    if (ll + nr - 1 < s || ll >= s + m) continue;
    const double *rec = p.mib.rec + (size_t)bi * kRecLen;
    double fa[5], fs[5];
    block_inputs<FLAGS>(p, ll, la, lb, fa, fs);
    double bv[3];
    for (int r = 0; r < 3; r++) bv[r] = r < nr ? block_row_rhs<FLAGS>(p, rec + r * kRecRow, fa, fs) : 0.0;
    // The right-hand side goes in as a select of register values; a select
    // between a global load and an array element would be turned into a load
    // through a selected pointer, which keeps the array in scratch memory.
    const R bv0 = (R)bv[0], bv1 = (R)bv[1], bv2 = (R)bv[2];
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      const int rr = s + r - ll;
      const bool hit = r < m && rr >= 0 && rr < nr;
      if (SOLVE && hit) {
        const double *lhs = rec + rr * kRecRow + kRecLhs;
        a[r] = (R)lhs[0];
        b[r] = (R)lhs[1];
        c[r] = (R)lhs[2];
      }
      const R dv = rr == 0 ? bv0 : (rr == 1 ? bv1 : bv2);
      d[r] = hit ? dv : d[r];
    }
  }
}

// Right-hand side of rows [s, s + m) of a regular line (no MIB blocks, every
// row on side `out`, whose coefficient is constant): the regular rows of
// line_rows, operation for operation, without the matrix coefficients, which
// the regular solve takes from its precomputed table.  The boundary rows'
// couplings are not applied here.
template <int DIR, int MMAX, int FLAGS, typename R = double>
MTS_HD void regular_rhs(const LineArgs &p, int j, int k, bool out, const R *la, const R *lb,
                        const double *wl, const LineSource<DIR> &src, const LaneSeg &sg,
                        R (&d)[MMAX]) {
  constexpr int RHS = FLAGS & 3;
  const LinePart &q = p.q;
  const int s = sg.s, m = sg.m;
#pragma unroll
  for (int r = 0; r < MMAX; r++) d[r] = 0.0;
  if (m == 0) return;
  const int base = 1 + sg.k * q.stride;
  const int sl_prev = sg.k == 0 ? 0 : base - q.stride + q.m - 1;
  const int sl_next = sg.k == q.p - 1 ? 1 + q.p * q.stride : base + q.stride;
  (void)src;
  const double *w = wl + (out ? 0 : 3);
  const R w0 = (R)w[0], w1 = (R)w[1], w2 = (R)w[2];
  const R rt = (R)p.theta_dt;
  (void)rt;
  (void)s;
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    if (r < m) {
      if (RHS == kRhsCopy) {
        d[r] = la[base + r];
      } else {
        const R *f = RHS == kRhsDouglas ? lb : la;
        const R fp = r == 0 ? f[sl_prev] : f[base + r - 1];
        const R fc = f[base + r];
        const R fn = r + 1 < m ? f[base + r + 1] : f[sl_next];
        const R sum = fmad(fn, w2, fmad(fc, w1, fp * w0));
        if (RHS == kRhsDouglas) {
          d[r] = fmad(-(R)p.dt, sum, la[base + r]);
        } else {
          R v = fmad(rt, sum, fc);
          if (RHS == kRhsTSZ) {
            v = v + (R)(p.c_f1 * src.at(p.P, p.g, p.e_f1, out, j, k, s + r));
            v = v + (R)(p.c_f2 * src.at(p.P, p.g, p.e_f2, out, j, k, s + r));
          }
          d[r] = v;
        }
      }
    }
  }
}

// Whether line lid is regular for the solve: no MIB blocks, and a constant
// diffusion coefficient on its side (outside always; inside when beta is
// constant).  Lines outside the interior (valid false) count as regular:
// they solve nothing.
MTS_HD bool regular_line(const PhysConst &P, const LineMeta &lm, bool valid) {
  return !valid || (lm.bb == lm.be && (lm.out0 || P.beta_kind == 0));
}

// Epilogue of the lane's rows into the LDS output tile; the group's lane 0
// also writes the two boundary rows (x0, xn: their values).
template <int DIR, int MMAX, int FLAGS, typename R = double>
MTS_HD void line_write(const LineArgs &p, int j, int k, const LineMeta &lm, const LaneSeg &sg,
                       const R (&x)[MMAX], double x0, double xn, R *lo,
                       const LineSource<DIR> &src) {
  constexpr bool EPI = (FLAGS & kFlagEpiF) != 0;
  using M = LineMap<DIR>;
  const LinePart &q = p.q;
  (void)src;
  if (sg.m > 0) {
    const unsigned cm = EPI ? rows_outside(p.mib, lm, sg.s) : 0u;
    const int base = 1 + sg.k * q.stride;
#pragma unroll
    for (int r = 0; r < MMAX; r++) {
      if (r < sg.m) {
        double v = x[r];
        if (EPI) v = v + p.c_f1 * src.at(p.P, p.g, p.e_f1, ((cm >> r) & 1u) != 0, j, k, sg.s + r);
        lo[base + r] = (R)v;
      }
    }
  }
  if (sg.k == 0 && sg.m > 0) {
    const int n = q.n;
    if (EPI) {
      int ix, iy, iz;
      M::node(j, k, 0, ix, iy, iz);
      x0 = x0 + p.c_f1 * source(p.P, p.g, p.e_f1, row_outside(p.mib, lm, 0), ix, iy, iz);
      M::node(j, k, n - 1, ix, iy, iz);
      xn = xn + p.c_f1 * source(p.P, p.g, p.e_f1, row_outside(p.mib, lm, n - 1), ix, iy, iz);
    }
    lo[0] = (R)x0;
    lo[1 + q.p * q.stride] = (R)xn;
  }
}

template <int DIR, int MMAX, int FLAGS, typename R = double>
MTS_HD void line_write(const LineArgs &p, int j, int k, const LineMeta &lm, const LaneSeg &sg,
                       const R (&x)[MMAX], double x0, double xn, R *lo, const double *zt) {
  const LineSource<DIR> src(p.P, p.g, j, k, zt);
  line_write<DIR, MMAX, FLAGS, R>(p, j, k, lm, sg, x, x0, xn, lo, src);
}

// Boundary values of line (j, k) for the output time (rows 0 and n-1).
template <int DIR>
MTS_HD void line_ends(const LineArgs &p, int j, int k, double &x0, double &xn) {
  using M = LineMap<DIR>;
  int ix, iy, iz;
  M::node(j, k, 0, ix, iy, iz);
  x0 = outer_u(p.g, p.e_out, ix, iy, iz);
  M::node(j, k, p.q.n - 1, ix, iy, iz);
  xn = outer_u(p.g, p.e_out, ix, iy, iz);
}

}  // namespace mts
