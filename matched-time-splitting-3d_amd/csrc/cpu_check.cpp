// CPU check of the per-step kernels: memory accesses and values.
//
// Builds the device table for one geometry exactly as an instance does, then
// runs one time step of the chosen method on the host, using the same
// per-node / per-lane functions the HIP kernels call (device_core.hpp), over
// buffers of exactly the device sizes.  Compiled with AddressSanitizer, so
// any index a kernel would form outside its buffer aborts here instead of
// faulting a GPU.  Every line stage is emulated end to end: the tile walk,
// the rows each lane builds (line_rows), the boundary couplings moved to the
// right-hand side as the kernel does, a sequential Thomas solve of the line
// on the host in place of the lane-parallel one (which touches registers
// only), and the write-back through line_write.  With --dump DIR the step
// starts from the exact initial condition at t = 0 with the time factors of
// the first step, and the fields after each stage are written to DIR for a
// comparison with the oracle.
//
//   mts_cpucheck <method adi|lod-ie|lod-cn|ts> <mib 1|2> <nodes per axis>
//                <surface ellipsoid|torus|bumpy> <p0> <p1> [p2 p3]
//                <beta const|sincos> <b0> <b1> [--dump DIR]
//
//   mts_cpucheck solve <nodes> <min lanes> <k>
//     the regular lines' right-hand-side-only solve (regular_solve.hpp) on a
//     line of <nodes> nodes with rows (-k, 1 + 2k, -k), every lane of the
//     group emulated phase by phase, against a Thomas solve of the same line;
//     prints the largest relative difference of the full and the guarded
//     variant.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "device_core.hpp"
#include "diagnostics.hpp"
#include "host_setup.hpp"
#include "host_table.hpp"
#include "regular_solve.hpp"
#include "stage_records.hpp"

using namespace mts;

namespace {

struct Emu {
  const SetupResult *s = nullptr;
  DeviceTableHost t;
  int px = 0;
  size_t fe = 0;
  std::vector<double> u, r;
  std::vector<int8_t> ind;
  DevGeom g;
  DevMib mib[3];
  PhysConst phys;
  long long reg_rows_checked = 0;  // regular rows whose right-hand sides were compared
};

// One pair of a thread's walk against the line and row it stands for:
// global index, LDS slots, existence and write-back recomputed directly.
template <int DIR, int T>
void check_pair(const LineArgs &p, const TileWalk<DIR, T> &w, const WalkPair &t, int j0, int k) {
  using M = LineMap<DIR>;
  const int n = p.g.n[DIR], nj = M::nj(p.g);
  const int j1 = DIR == 0 ? w.j : w.j + 1;      // line of the second element
  const int r1 = DIR == 0 ? t.r + 1 : t.r;      // row of the second element
  const bool ok0 = w.j < nj && t.r < n, ok1 = j1 < nj && r1 < n;
  if (t.ok0 != ok0 || t.ok1 != ok1) std::abort();
  if (t.w0 != (ok0 && w.j >= 1 && w.j <= nj - 2)) std::abort();
  if (t.w1 != (ok1 && j1 >= 1 && j1 <= nj - 2)) std::abort();
  const bool end0 = t.r == 0 || t.r == n - 1, end1 = r1 == 0 || r1 == n - 1;
  if (t.ld0 != (ok0 && !(p.in_analytic && end0))) std::abort();
  if (t.ld1 != (ok1 && !(p.in_analytic && end1))) std::abort();
  if (ok0) {
    if (t.gi != M::base(p.g, w.j, k) + (size_t)t.r * M::stride(p.g)) std::abort();
    if (t.gi % 2 != 0) std::abort();  // 16-byte alignment of the pair
    if (t.sl0 != (w.j - j0) * p.q.lds_len + lds_slot(p.q, t.r)) std::abort();
  }
  if (ok1) {
    if (t.gi + 1 != M::base(p.g, j1, k) + (size_t)r1 * M::stride(p.g)) std::abort();
    if (t.sl1 != (j1 - j0) * p.q.lds_len + lds_slot(p.q, r1)) std::abort();
  }
}

template <int DIR, int L, int MMAX, int FLAGS>
void emulate_line(Emu &E, const LineArgs &p) {
  constexpr int RHS = FLAGS & 3;
  constexpr bool SOLVE = (FLAGS & kFlagSolve) != 0;
  constexpr bool HAS_S = (RHS == kRhsDouglas);
  constexpr bool TWO = HAS_S || !SOLVE;
  constexpr bool PRE1 = DIR == 0 && (FLAGS & kFlagPre1) != 0;
  constexpr int T = tile_lines(L);
  const LinePart &q = p.q;
  const int n = p.g.n[DIR];
  const int nj = LineMap<DIR>::nj(p.g), nk = LineMap<DIR>::nk(p.g);
  const int ntile = (nj + T - 1) / T;
  for (int bx = 0; bx < ntile; bx++)
    for (int k = 1; k <= nk - 2; k++) {
      // exact-size LDS tiles
      std::unique_ptr<double[]> tA(new double[(size_t)T * q.lds_len]);
      std::unique_ptr<double[]> tB(new double[(size_t)T * q.lds_len]);
      std::memset(tA.get(), 0, sizeof(double) * T * q.lds_len);
      std::memset(tB.get(), 0, sizeof(double) * T * q.lds_len);
      const int j0 = bx * T;
      if (T * n > kTileElsPerThread * kBlockThreads) std::abort();
      // the kernel's paired 16-byte loads: both doubles are read whenever the
      // first element exists; every element of the tile is moved exactly once
      std::vector<int> seen((size_t)T * q.lds_len, 0);
      for (int tid = 0; tid < kBlockThreads; tid++) {
        const TileWalk<DIR, T> w(p, j0, k, tid);
        for (int u = 0; u < kTileSteps; u++) {
          const WalkPair t = w.at(p, u);
          check_pair<DIR, T>(p, w, t, j0, k);
          if (PRE1) {
            // the x sweep's tile of u^n staged as the stage-1 right-hand side
            // (interior lines only), pair by pair as the kernel does
            if (!(t.ld0 && t.w0) && !(t.ld1 && t.w1)) continue;
            Pre1Pair e;
            pre1_load(p, t.gi, t.r, e);
            const Pre1Line ln = pre1_line(p.g, w.j, k);
            if (t.ld0 && t.w0) {
              tA[t.sl0] = pre1_value(p, ln, e, t.gi, 0);
              seen[t.sl0]++;
            }
            if (t.ld1 && t.w1) {
              tA[t.sl1] = pre1_value(p, ln, e, t.gi, 1);
              seen[t.sl1]++;
            }
            continue;
          }
          if (!t.ld0 && !t.ld1) continue;
          const double *fA = static_cast<const double *>(p.A);
          const double *fS = static_cast<const double *>(p.S);
          const double a0 = fA[t.gi], a1 = fA[t.gi + 1];
          double s0 = 0.0, s1 = 0.0;
          if (HAS_S) {
            s0 = fS[t.gi];
            s1 = fS[t.gi + 1];
          }
          if (t.ld0) {
            tA[t.sl0] = a0;
            if (HAS_S) tB[t.sl0] = s0;
            seen[t.sl0]++;
          }
          if (t.ld1) {
            tA[t.sl1] = a1;
            if (HAS_S) tB[t.sl1] = s1;
            seen[t.sl1]++;
          }
        }
      }
      // in_analytic stages: the end rows come from tile_end_val, not the walk
      if (p.in_analytic)
        for (int e = 0; e < 2 * T; e++) {
          const int sl = tile_end_slot<DIR>(p, j0, e);
          if (sl < 0) continue;
          tile_end_value<DIR>(p, j0, k, e, tA.get());
          seen[sl]++;
        }
      for (int l = 0; l < T && j0 + l < nj; l++) {
        if (PRE1 && (j0 + l < 1 || j0 + l > nj - 2)) continue;  // (not staged)
        for (int r = 0; r < n; r++)
          if (seen[(size_t)l * q.lds_len + lds_slot(q, r)] != 1) std::abort();
      }
      std::unique_ptr<double[]> wt(new double[(size_t)T * kWtLen]);
      for (int e = 0; e < T * kWtLen; e++) line_weight_entry<DIR>(p, j0, k, e, wt.get());
      std::vector<double> zt(2 * (size_t)p.g.n[2]);
      for (int e = 0; e < 2 * p.g.n[2]; e++) ztab_entry(p.g, e, zt.data());
      for (int l = 0; l < T; l++) {
        const int j = j0 + l;
        const bool valid = j >= 1 && j <= nj - 2;
        const LineMeta lm = line_meta(p.mib, k * nj + j, valid);
        // the rows of every lane of the line's group, then the line solved as
        // the kernel solves it (boundary couplings on the right-hand side)
        std::vector<double> la(L * MMAX), lb(L * MMAX), lc(L * MMAX), ld(L * MMAX);
        std::vector<LaneSeg> segs(L);
        for (int kseg = 0; kseg < L; kseg++) {
          const LaneSeg sg = lane_seg(q, kseg, valid);
          segs[kseg] = sg;
          if (sg.m > 0 && (sg.m < 2 || sg.m > MMAX || sg.s < 1 || sg.s + sg.m > n - 1)) std::abort();
          double a[MMAX], b[MMAX], c[MMAX], d[MMAX];
          line_rows<DIR, MMAX, FLAGS>(p, j, k, lm, tA.get() + l * q.lds_len,
                                      tB.get() + l * q.lds_len, wt.get() + l * kWtLen,
                                      DIR == 2 ? zt.data() : nullptr, sg, a, b, c, d);
          for (int r = 0; r < MMAX; r++) {
            la[kseg * MMAX + r] = a[r];
            lb[kseg * MMAX + r] = b[r];
            lc[kseg * MMAX + r] = c[r];
            ld[kseg * MMAX + r] = d[r];
          }
          if (SOLVE && regular_line(p.P, lm, valid)) {
            // the regular solve's right-hand side is the general one, bit for bit
            double dr[MMAX];
            const LineSource<DIR> lsrc(p.P, p.g, j, k, DIR == 2 ? zt.data() : nullptr);
            regular_rhs<DIR, MMAX, FLAGS>(p, j, k, lm.out0, tA.get() + l * q.lds_len,
                                          tB.get() + l * q.lds_len, wt.get() + l * kWtLen, lsrc,
                                          sg, dr);
            if (std::memcmp(dr, d, sizeof(dr)) != 0) std::abort();
            E.reg_rows_checked += sg.m;
          }
        }
        if (!valid) continue;
        double x0 = 0.0, xn = 0.0;
        line_ends<DIR>(p, j, k, x0, xn);
        if (SOLVE) {
          // interior rows 1 .. n-2 in line order
          std::vector<double> A(n), B(n), C(n), D(n), X(n);
          std::vector<int> seen_row(n, 0);
          for (int kseg = 0; kseg < L; kseg++)
            for (int r = 0; r < segs[kseg].m; r++) {
              const int i = segs[kseg].s + r;
              A[i] = la[kseg * MMAX + r];
              B[i] = lb[kseg * MMAX + r];
              C[i] = lc[kseg * MMAX + r];
              D[i] = ld[kseg * MMAX + r];
              seen_row[i]++;
            }
          for (int i = 1; i <= n - 2; i++)
            if (seen_row[i] != 1) std::abort();  // the segments cover the interior exactly once
          D[1] -= A[1] * x0;
          D[n - 2] -= C[n - 2] * xn;
          // Thomas over rows 1 .. n-2 (the reference's TDMA order)
          std::vector<double> gam(n);
          double bet = B[1];
          X[1] = D[1] / bet;
          for (int i = 2; i <= n - 2; i++) {
            gam[i] = C[i - 1] / bet;
            bet = B[i] - A[i] * gam[i];
            X[i] = (D[i] - A[i] * X[i - 1]) / bet;
          }
          for (int i = n - 3; i >= 1; i--) X[i] -= gam[i + 1] * X[i + 1];
          for (int kseg = 0; kseg < L; kseg++)
            for (int r = 0; r < segs[kseg].m; r++) ld[kseg * MMAX + r] = X[segs[kseg].s + r];
        }
        for (int kseg = 0; kseg < L; kseg++) {
          double x[MMAX];
          for (int r = 0; r < MMAX; r++) x[r] = ld[kseg * MMAX + r];
          line_write<DIR, MMAX, FLAGS>(p, j, k, lm, segs[kseg], x, x0, xn,
                                       (TWO && !HAS_S ? tB.get() : tA.get()) + l * q.lds_len,
                                       DIR == 2 ? zt.data() : nullptr);
        }
      }
      const double *tO = TWO && !HAS_S ? tB.get() : tA.get();
      for (int tid = 0; tid < kBlockThreads; tid++) {
        const TileWalk<DIR, T> w(p, j0, k, tid);
        for (int u = 0; u < kTileSteps; u++) {
          const WalkPair t = w.at(p, u);
          const double v0 = t.w0 ? tO[t.sl0] : 0.0, v1 = t.w1 ? tO[t.sl1] : 0.0;
          if (t.w0) static_cast<double *>(p.out)[t.gi] = v0;
          if (t.w1) static_cast<double *>(p.out)[t.gi + 1] = v1;
        }
      }
    }
}

template <int FLAGS>
void line(Emu &E, int dir, const LineArgs &args) {
  LineArgs p = args;
  if (!make_line_part(p.g.n[dir], p.q, p.min_lanes)) std::abort();
  std::vector<double> rec;
  const std::string rerr =
      build_stage_records(*E.s, E.t, dir, FLAGS, p.theta_dt, p.dt, 0.5 * p.dt, 0.5 * p.dt, rec);
  if (!rerr.empty()) std::abort();
  p.mib.rec = rec.data();
  auto run = [&](auto lc, auto mc) {
    constexpr int LL = decltype(lc)::value;
    constexpr int MM = decltype(mc)::value;
    if (dir == 0) emulate_line<0, LL, MM, FLAGS>(E, p);
    else if (dir == 1) emulate_line<1, LL, MM, FLAGS>(E, p);
    else emulate_line<2, LL, MM, FLAGS>(E, p);
  };
  using I16 = std::integral_constant<int, 16>;
  using I32 = std::integral_constant<int, 32>;
  using I64 = std::integral_constant<int, 64>;
  using M4 = std::integral_constant<int, 4>;
  using M8 = std::integral_constant<int, 8>;
  using M16 = std::integral_constant<int, 16>;
  if (p.q.L == 16 && p.q.mmax == 4) run(I16(), M4());
  else if (p.q.L == 16 && p.q.mmax == 8) run(I16(), M8());
  else if (p.q.L == 16 && p.q.mmax == 16) run(I16(), M16());
  else if (p.q.L == 32 && p.q.mmax == 4) run(I32(), M4());
  else if (p.q.L == 32 && p.q.mmax == 8) run(I32(), M8());
  else if (p.q.L == 32 && p.q.mmax == 16) run(I32(), M16());
  else if (p.q.L == 64 && p.q.mmax == 4) run(I64(), M4());
  else if (p.q.L == 64 && p.q.mmax == 8) run(I64(), M8());
  else if (p.q.L == 64 && p.q.mmax == 16) run(I64(), M16());
  else std::abort();
}

// The regular lines' right-hand-side-only solve of one line, every lane of
// the group emulated: the per-lane phases are the kernels' own functions, the
// cross-lane shifts are done here.  Returns the largest difference from x.
template <int MMAX, bool FULL>
double regular_line(const LinePart &q, const std::vector<double> &tab, const std::vector<double> &d,
                    const std::vector<double> &x) {
  const int L = q.L, n = q.n;
  std::vector<std::array<double, MMAX>> dl((size_t)L);
  std::vector<double> D1((size_t)L), D2((size_t)L), ed((size_t)L), xe((size_t)L);
  std::vector<int> ms((size_t)L);
  for (int k = 0; k < L; k++) {
    const LaneSeg sg = lane_seg(q, k, true);
    ms[(size_t)k] = sg.m;
    double row[MMAX];
    for (int r = 0; r < MMAX; r++) row[r] = r < sg.m ? d[(size_t)(sg.s + r)] : 0.0;
    const double *rows = tab.data() + reg_rows_off(k == q.p - 1 ? 1 : 0, MMAX);
    const double *sc = tab.data() + reg_scal_off(k, MMAX);
    reg_forward<MMAX, FULL>(rows, sc, row, sg.m, D1[(size_t)k], D2[(size_t)k]);
    for (int r = 0; r < MMAX; r++) dl[(size_t)k][(size_t)r] = row[r];
  }
  for (int k = 0; k < L; k++) {
    const double *sc = tab.data() + reg_scal_off(k, MMAX);
    const double D1n = k + 1 < L ? D1[(size_t)k + 1] : 0.0;
    ed[(size_t)k] = reg_fold(sc, FULL ? MMAX : ms[(size_t)k], D1[(size_t)k], D2[(size_t)k], D1n);
  }
  for (int lv = 0, s = 1; s < L; lv++, s <<= 1) {
    std::vector<double> nd((size_t)L);
    for (int k = 0; k < L; k++) {
      const double *sc = tab.data() + reg_scal_off(k, MMAX);
      const double dm = k >= s ? ed[(size_t)(k - s)] : 0.0;
      const double dp = k + s < L ? ed[(size_t)(k + s)] : 0.0;
      nd[(size_t)k] = reg_level(sc, lv, ed[(size_t)k], dm, dp);
    }
    ed = nd;
  }
  for (int k = 0; k < L; k++) xe[(size_t)k] = ed[(size_t)k] * tab[(size_t)reg_scal_off(k, MMAX) + kRegQ];
  double err = 0.0, mx = 0.0;
  for (int k = 0; k < L; k++) {
    const LaneSeg sg = lane_seg(q, k, true);
    const double *rows = tab.data() + reg_rows_off(k == q.p - 1 ? 1 : 0, MMAX);
    const double *sc = tab.data() + reg_scal_off(k, MMAX);
    double row[MMAX];
    for (int r = 0; r < MMAX; r++) row[r] = dl[(size_t)k][(size_t)r];
    const double xprev = k >= 1 ? xe[(size_t)k - 1] : 0.0;
    reg_back<MMAX, FULL>(rows, sc, row, sg.m, D1[(size_t)k], xe[(size_t)k], xprev);
    for (int r = 0; r < sg.m; r++) {
      err = std::max(err, std::fabs(row[r] - x[(size_t)(sg.s + r)]));
      mx = std::max(mx, std::fabs(x[(size_t)(sg.s + r)]));
    }
  }
  (void)n;
  return err / mx;
}

// The Douglas stage-1 right-hand side at one node written with each load
// where its value is used (the kernel's node function loads everything up
// front); the two must agree bit for bit.
double rhs1_plain(const Rhs1Args &p, int ix, int iy, int iz) {
  const DevGeom &g = p.g;
  const size_t c = node_index(g, ix, iy, iz);
  if (ix == 0 || iy == 0 || iz == 0 || ix == g.n[0] - 1 || iy == g.n[1] - 1 || iz == g.n[2] - 1)
    return outer_u(g, p.e, ix, iy, iz);
  const bool out = g.ind[c] > 0;
  const size_t sy = (size_t)g.px, sz = (size_t)g.n[1] * g.px;
  double vy[3], vz[3];
  op_weights(p.P, g, 1, out, ix, iy, iz, vy);
  op_weights(p.P, g, 2, out, ix, iy, iz, vz);
  const double *pu = static_cast<const double *>(p.u);
  const double u = pu[c];
  double sum_y = pu[c - sy] * vy[0] + u * vy[1] + pu[c + sy] * vy[2];
  double sum_z = pu[c - sz] * vz[0] + u * vz[1] + pu[c + sz] * vz[2];
  double r = p.dt * sum_y;
  r += p.dt * sum_z;
  r += u;
  r += p.dt * source(p.P, g, p.e, out, ix, iy, iz);
  return r;
}

int check_solve(int n, int min_lanes, double kk) {
  LinePart q;
  if (!make_line_part(n, q, min_lanes)) return 2;
  const double a = -kk, b = 1.0 + 2.0 * kk, c = -kk;
  // a right-hand side on rows 1 .. n-2 with the boundary couplings moved in
  std::vector<double> d((size_t)n, 0.0), x((size_t)n, 0.0);
  unsigned s = 12345u + (unsigned)n;
  for (int i = 1; i <= n - 2; i++) {
    s = s * 1103515245u + 12345u;
    d[(size_t)i] = ((s >> 8) & 0xffff) / 65536.0 - 0.5;
  }
  // Thomas over rows 1 .. n-2
  {
    std::vector<double> gam((size_t)n);
    double bet = b;
    x[1] = d[1] / bet;
    for (int i = 2; i <= n - 2; i++) {
      gam[(size_t)i] = c / bet;
      bet = b - a * gam[(size_t)i];
      x[(size_t)i] = (d[(size_t)i] - a * x[(size_t)i - 1]) / bet;
    }
    for (int i = n - 3; i >= 1; i--) x[(size_t)i] -= gam[(size_t)i + 1] * x[(size_t)i + 1];
  }
  double worst[2] = {-1.0, -1.0};
  for (int f = 0; f < 2; f++) {
    const bool full = f == 1;
    if (full && q.m != q.mmax) continue;
    std::vector<double> tab;
    if (!regular_factors(q, q.mmax, full, a, b, c, tab)) return 4;
    double e = 0.0;
    if (q.mmax == 4) e = full ? regular_line<4, true>(q, tab, d, x) : regular_line<4, false>(q, tab, d, x);
    else if (q.mmax == 8) e = full ? regular_line<8, true>(q, tab, d, x) : regular_line<8, false>(q, tab, d, x);
    else e = full ? regular_line<16, true>(q, tab, d, x) : regular_line<16, false>(q, tab, d, x);
    worst[f] = e;
  }
  std::printf("ok solve n %d L %d mmax %d m %d mlast %d p %d guarded %.3e full %.3e\n", n, q.L,
              q.mmax, q.m, q.mlast, q.p, worst[0], worst[1]);
  return 0;
}

// Field (x-pitched) to nx*ny*nz, x fastest, written as raw float64.
void dump_field(const std::string &path, const std::vector<double> &f, const int nn[3], int px) {
  std::vector<double> out((size_t)nn[0] * nn[1] * nn[2]);
  for (int iz = 0; iz < nn[2]; iz++)
    for (int iy = 0; iy < nn[1]; iy++)
      std::memcpy(&out[((size_t)iz * nn[1] + iy) * nn[0]], &f[((size_t)iz * nn[1] + iy) * px],
                  nn[0] * sizeof(double));
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp || std::fwrite(out.data(), sizeof(double), out.size(), fp) != out.size()) std::abort();
  std::fclose(fp);
}

// Boundary faces of one step (faces_kernel).
void faces(const Emu &E, std::vector<double> &u, double e, double cf) {
  for (int iz = 0; iz < E.g.n[2]; iz++)
    for (int iy = 0; iy < E.g.n[1]; iy++)
      for (int ix = 0; ix < E.g.n[0]; ix++) {
        const bool face = ix == 0 || iy == 0 || iz == 0 || ix == E.g.n[0] - 1 ||
                          iy == E.g.n[1] - 1 || iz == E.g.n[2] - 1;
        if (face) u[node_index(E.g, ix, iy, iz)] = face_value(E.phys, E.g, e, cf, ix, iy, iz);
      }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc == 5 && std::string(argv[1]) == "solve")
    return check_solve(std::atoi(argv[2]), std::atoi(argv[3]), std::atof(argv[4]));
  std::string dump_dir;
  if (argc > 2 && std::string(argv[argc - 2]) == "--dump") {
    dump_dir = argv[argc - 1];
    argc -= 2;
  }
  if (argc < 9) {
    std::fprintf(stderr, "usage: mts_cpucheck method mib nodes surface p.. beta b0 b1 [--dump DIR]\n");
    return 2;
  }
  std::string method = argv[1];
  int mib = std::atoi(argv[2]);
  // nodes per axis: "N" or "NXxNYxNZ"
  int nn[3];
  if (std::sscanf(argv[3], "%dx%dx%d", &nn[0], &nn[1], &nn[2]) != 3) nn[0] = nn[1] = nn[2] = std::atoi(argv[3]);
  std::string sk = argv[4];
  int ai = 5;
  std::unique_ptr<Surface> surf;
  if (sk == "ellipsoid") {
    surf.reset(new SurfaceEllipsoid(atof(argv[5]), atof(argv[6]), atof(argv[7]), atof(argv[8])));
    ai = 9;
  } else if (sk == "torus") {
    surf.reset(new SurfaceTorus(atof(argv[5]), atof(argv[6])));
    ai = 7;
  } else if (sk == "bumpy") {
    surf.reset(new SurfaceBumpy(atof(argv[5]), atof(argv[6]), atof(argv[7]), atof(argv[8])));
    ai = 9;
  } else {
    return 2;
  }
  BetaModel beta;
  beta.kind = std::string(argv[ai]) == "sincos" ? 1 : 0;
  beta.p[0] = atof(argv[ai + 1]);
  beta.p[1] = atof(argv[ai + 2]);
  double lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};
  Grid g = make_grid(lo, hi, nn);
  SetupResult res;
  BuildOptions opt;
  opt.mib = mib;
  std::string err = build_setup(g, *surf, beta, opt, res);
  if (!err.empty()) {
    std::printf("setup refused: %s\n", err.c_str());
    return 3;
  }
  const double dt = 2.0 / (nn[0] - 1);
  const bool adi = method == "adi";
  const double tdt = (adi || method == "lod-ie") ? dt : 0.5 * dt;
  Emu E;
  E.s = &res;
  const std::string terr = build_device_table(res, beta, tdt, adi, E.t);
  if (!terr.empty()) {
    std::printf("table refused: %s\n", terr.c_str());
    return 3;
  }
  E.px = (nn[0] + 15) / 16 * 16;
  E.fe = (size_t)E.px * nn[1] * nn[2];
  E.u.assign(E.fe, 0.25);
  E.r.assign(E.fe, 0.0);
  E.ind.assign(E.fe, 1);
  for (int iz = 0; iz < nn[2]; iz++)
    for (int iy = 0; iy < nn[1]; iy++)
      std::memcpy(&E.ind[((size_t)iz * nn[1] + iy) * E.px], &res.ind[g.idx(0, iy, iz)], nn[0]);
  for (int a = 0; a < 3; a++) {
    E.g.n[a] = nn[a];
    E.g.tab[a] = E.t.tab[a].data();
  }
  E.g.px = E.px;
  E.g.ind = E.ind.data();
  E.phys = E.t.phys;
  for (int d = 0; d < 3; d++) {
    // a direction without blocks holds no table (the device buffers of size
    // zero are null pointers): the kernels must not read it
    const bool any = E.t.n_blocks[d] > 0;
    E.mib[d].blk_ptr = E.t.blk_ptr[d].data();
    E.mib[d].hdr = any ? E.t.blk_hdr[d].data() : nullptr;
    E.mib[d].val = any ? E.t.blk_val[d].data() : nullptr;
    E.mib[d].cr = E.t.cr_val[d].empty() ? nullptr : E.t.cr_val[d].data();
    E.mib[d].seq = E.t.n_seq[d] > 0 ? E.t.seq[d].data() : nullptr;
    E.mib[d].cls0 = E.t.cls0[d].data();
  }
  // the solve stages' tile lists: every interior line in exactly one tile,
  // the regular tiles holding regular lines only
  {
    int ml = 16;
    if (const char *ev = std::getenv("MTS_LINE_LANES")) ml = std::atoi(ev);
    for (int d = 0; d < 3; d++) {
      const int nj = d == 0 ? nn[1] : nn[0], nk = d == 2 ? nn[1] : nn[2];
      LinePart qr, qg;
      make_line_part(nn[d], qr, ml);
      make_line_part(nn[d], qg, std::max(ml, 32));
      for (int eb : {8, 4}) {  // the lists of the FP64 and the FP32 form
      std::vector<int> lr, lg;
      if (!stage_tiles(E.t.blk_ptr[d].data(), E.t.cls0[d].data(), E.t.phys.beta_kind == 0, d,
                       nn[d], nj, nk, ml, 32, eb, lr, lg))
        std::abort();
      // grouped by line: in every full window of 8 line groups, the tiles
      // of one line group are 8 entries apart (same XCD)
      for (int c = 0; c < 2; c++) {
        const std::vector<int> &lst = c == 0 ? lr : lg;
        const int G = line_group(d, tile_lines(c == 0 ? qr.L : qg.L), eb);
        if (G <= 1) continue;
        // full line groups in the list (all G tiles of a line in it)
        const int ng = (nj + G - 1) / G + 1;
        std::vector<int> cnt((size_t)nk * ng, 0);
        for (int e : lst)
          if (tile_k(e) < nk && tile_bx(e) / G < ng) cnt[(size_t)tile_k(e) * ng + tile_bx(e) / G]++;
        size_t full = 0;
        for (int v : cnt) full += v == G ? 1 : 0;
        size_t i = 0, windows = 0;
        while (i + 8 * (size_t)G <= lst.size()) {
          bool window = true;  // a window of 8 full line groups?
          for (int u = 0; u < 8 && window; u++)
            for (int m = 1; m < G; m++) {
              const int a = lst[i + u], b = lst[i + (size_t)m * 8 + u];
              if (tile_k(a) != tile_k(b) || tile_bx(a) / G != tile_bx(b) / G) window = false;
            }
          if (!window) break;  // the rest: partial line groups and the tail
          i += 8 * (size_t)G;
          windows++;
        }
        if (windows != full / 8) {
          std::printf("dir %d: tile list not grouped by line (%zu windows, %zu full groups)\n", d,
                      windows, full);
          return 1;
        }
      }
      std::vector<int> cover((size_t)nj * nk, 0);
      for (int c = 0; c < 2; c++) {
        const int T = tile_lines(c == 0 ? qr.L : qg.L);
        for (int e : c == 0 ? lr : lg) {
          const int bx = tile_bx(e), k = tile_k(e);
          if (k < 1 || k > nk - 2 || bx * T >= nj) std::abort();
          for (int j = bx * T; j < bx * T + T && j < nj; j++) {
            if (j < 1 || j > nj - 2) continue;
            const int lid = k * nj + j;
            cover[(size_t)lid]++;
            const bool reg = E.t.blk_ptr[d][lid] == E.t.blk_ptr[d][lid + 1] &&
                             (E.t.cls0[d][lid] || E.t.phys.beta_kind == 0);
            if (c == 0 && !reg) {
              std::printf("dir %d: irregular line %d in a regular tile\n", d, lid);
              return 1;
            }
          }
        }
      }
      for (int k = 1; k <= nk - 2; k++)
        for (int j = 1; j <= nj - 2; j++)
          if (cover[(size_t)k * nj + j] != 1) {
            std::printf("dir %d: line (%d, %d) in %d tiles\n", d, j, k, cover[(size_t)k * nj + j]);
            return 1;
          }
      std::printf("dir %d tiles (%d-byte elements): %zu regular (L %d), %zu general (L %d)\n", d,
                  eb, lr.size(), qr.L, lg.size(), qg.L);
      }
    }
  }
  // the device interface diagnostics (diag_line, one call per line) against
  // the host diagnostics: equal to 1e-12 (NaN where the host has NaN).  Not
  // bit for bit: the device reads sin and cos of a crossing from its jump
  // record, and a compiler may form that record's sin and cos of one
  // argument with one sincos call, which can differ from sin alone by an ulp
  // (g++ does here)
  for (int d = 0; d < 3; d++) {
    const double t = 0.37;
    std::vector<double> hl, hr, hj;
    crossing_errors(res, beta, d, t, hl, hr, hj);
    const size_t nc = hl.size();
    std::vector<double> el(nc, -1.0), er(nc, -1.0), je(nc, -1.0);
    DiagArgs a;
    a.g = E.g;
    a.mib = E.mib[d];
    a.cr = E.t.cr_val[d].empty() ? nullptr : E.t.cr_val[d].data();
    a.nj = d == 0 ? nn[1] : nn[0];
    a.nk = d == 2 ? nn[1] : nn[2];
    a.e = std::exp(-t);
    a.el = el.data();
    a.er = er.data();
    a.je = je.data();
    for (int lid = 0; lid < a.nj * a.nk; lid++) {
      if (d == 0) diag_line<0>(a, lid);
      else if (d == 1) diag_line<1>(a, lid);
      else diag_line<2>(a, lid);
    }
    auto same = [](double x, double y) {
      if (std::isnan(x) || std::isnan(y)) return std::isnan(x) && std::isnan(y);
      if (std::isinf(x) || std::isinf(y)) return x == y;
      return std::fabs(x - y) <= 1e-12;
    };
    for (size_t c = 0; c < nc; c++)
      if (!same(el[c], hl[c]) || !same(er[c], hr[c]) || !same(je[c], hj[c])) {
        std::printf("dir %d crossing %zu: device diagnostics differ from the host's\n", d, c);
        std::printf("%.17g %.17g | %.17g %.17g | %.17g %.17g id %d ll %d\n", el[c], hl[c], er[c], hr[c], je[c], hj[c], res.cross[d][c].id, res.cross[d][c].left_loc);
        return 1;
      }
  }
  // time factors: the first step from t = 0 (mts_step with t_next = dt)
  const double t1 = dt;
  const double e1 = std::exp(-t1), en = std::exp(-(t1 - dt)), eh = std::exp(-(t1 - 0.5 * dt));
  const bool dump = !dump_dir.empty();
  if (dump) {
    for (int iz = 0; iz < nn[2]; iz++)
      for (int iy = 0; iy < nn[1]; iy++)
        for (int ix = 0; ix < nn[0]; ix++) {
          const size_t c = node_index(E.g, ix, iy, iz);
          E.u[c] = E.ind[c] > 0 ? outer_u(E.g, 1.0, ix, iy, iz) : inner_u(E.g, 1.0, ix, iy, iz);
        }
    dump_field(dump_dir + "/u0.bin", E.u, nn, E.px);
  }
  LineArgs a;
  std::memset(&a, 0, sizeof(a));
  a.min_lanes = 16;
  if (const char *ev = std::getenv("MTS_LINE_LANES")) a.min_lanes = std::atoi(ev);
  a.g = E.g;
  a.P = E.t.phys;
  a.theta_dt = tdt;
  a.dt = dt;
  if (adi) {
    // the stage-1 launch: every (x row, node pair) the pair threads cover,
    // the MIB-affected nodes (classification +-2, as the instance plan marks
    // them) left to spec_node; the pair form must equal the plain per-node
    // form bit for bit, and the spec nodes must be exactly the ones it skips
    for (size_t r = 0; r < E.t.spec.size(); r += kSpecLen) {
      int8_t &v = E.ind[((size_t)E.t.spec[r + 2] * nn[1] + E.t.spec[r + 1]) * E.px + E.t.spec[r]];
      v = v > 0 ? 2 : -2;
    }
    Rhs1Args r1{E.g, E.t.phys, E.u.data(), E.r.data(), dt, e1};
    const double kUnset = -12345.5;
    std::fill(E.r.begin(), E.r.end(), kUnset);
    const unsigned hp = (unsigned)E.px / 2;
    for (unsigned row = 0; row < (unsigned)(nn[1] * nn[2]); row++)
      for (unsigned pp = 0; pp < hp; pp++) rhs1_pair(r1, row, pp);
    std::vector<double> plain(E.fe, 0.0);
    for (int iz = 0; iz < nn[2]; iz++)
      for (int iy = 0; iy < nn[1]; iy++)
        for (int ix = 0; ix < nn[0]; ix++) {
          const size_t c = node_index(E.g, ix, iy, iz);
          const bool spec = E.ind[c] == 2 || E.ind[c] == -2;
          const double v = rhs1_plain(r1, ix, iy, iz);
          if (spec ? E.r[c] != kUnset : std::memcmp(&v, &E.r[c], sizeof v) != 0) {
            std::printf("rhs1 pair at node (%d, %d, %d) differs from the plain form\n", ix, iy,
                        iz);
            return 1;
          }
          rhs1_node(r1, ix, iy, iz);
          if (std::memcmp(&v, &E.r[c], sizeof v) != 0) {
            std::printf("rhs1 node (%d, %d, %d) differs from the plain form\n", ix, iy, iz);
            return 1;
          }
        }
    SpecArgs sp;
    sp.g = E.g;
    sp.P = E.t.phys;
    for (int d = 0; d < 3; d++) sp.mib[d] = E.mib[d];
    sp.spec = E.t.spec.data();
    sp.nspec = (int)(E.t.spec.size() / kSpecLen);
    sp.u = E.u.data();
    sp.R = E.r.data();
    sp.dt = dt;
    sp.e = e1;
    for (int t = 0; t < sp.nspec; t++) spec_node(sp, t);
    if (dump) dump_field(dump_dir + "/rhs1.bin", E.r, nn, E.px);
    // the fused x sweep (the default): r holds only the MIB-affected nodes'
    // values, the sweep stages u^n as R; its output must equal the separate
    // stage-1 pass followed by the plain x sweep, bit for bit
    std::vector<double> r_fused(E.fe, -7.5);
    for (size_t r = 0; r < E.t.spec.size(); r += kSpecLen) {
      const size_t c = node_index(E.g, E.t.spec[r], E.t.spec[r + 1], E.t.spec[r + 2]);
      r_fused[c] = E.r[c];
    }
    a.e_out = e1;
    a.e_jump = e1;
    a.mib = E.mib[0];
    {
      LineArgs af = a;
      af.A = E.u.data();
      af.S = r_fused.data();
      af.out = r_fused.data();
      af.in_analytic = 1;
      af.e_in = e1;
      line<kRhsCopy | kFlagSolve | kFlagPre1>(E, 0, af);
    }
    a.A = E.r.data();
    a.out = E.r.data();
    line<kRhsCopy | kFlagSolve>(E, 0, a);
    for (int iz = 1; iz < nn[2] - 1; iz++)
      for (int iy = 1; iy < nn[1] - 1; iy++)
        for (int ix = 0; ix < nn[0]; ix++) {
          const size_t c = node_index(E.g, ix, iy, iz);
          if (std::memcmp(&r_fused[c], &E.r[c], sizeof(double)) != 0) {
            std::printf("fused x sweep differs at node (%d, %d, %d): %.17g vs %.17g\n", ix, iy,
                        iz, r_fused[c], E.r[c]);
            return 1;
          }
        }
    if (dump) dump_field(dump_dir + "/ustar.bin", E.r, nn, E.px);
    a.S = E.u.data();
    a.mib = E.mib[1];
    line<kRhsDouglas | kFlagSolve>(E, 1, a);
    if (dump) dump_field(dump_dir + "/ustar2.bin", E.r, nn, E.px);
    a.mib = E.mib[2];
    a.out = E.u.data();
    line<kRhsDouglas | kFlagSolve>(E, 2, a);
    faces(E, E.u, e1, 0.0);
  } else if (method == "lod-ie" || method == "lod-cn") {
    a.e_out = e1;
    a.e_jump = e1;
    a.A = a.out = E.u.data();
    const bool cn = method == "lod-cn";
    for (int d = 0; d < 3; d++) {
      a.mib = E.mib[d];
      if (d == 1) {
        a.in_analytic = 1;
        a.e_in = e1;
      }
      if (d == 2) {
        a.e_f1 = e1;
        a.c_f1 = dt;
      }
      if (cn) {
        if (d < 2) line<kRhsCN | kFlagSolve | kFlagLhsJump>(E, d, a);
        else line<kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF>(E, d, a);
      } else {
        if (d < 2) line<kRhsCopy | kFlagSolve | kFlagLhsJump>(E, d, a);
        else line<kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF>(E, d, a);
      }
    }
    faces(E, E.u, e1, dt);
  } else if (method == "ts") {
    a.e_jump = e1;
    a.A = a.out = E.u.data();
    a.e_out = eh;
    a.mib = E.mib[0];
    line<kRhsCN>(E, 0, a);
    a.in_analytic = 1;
    a.e_in = eh;
    a.mib = E.mib[1];
    line<kRhsCN>(E, 1, a);
    a.mib = E.mib[2];
    a.e_out = e1;
    a.e_f1 = en;
    a.c_f1 = 0.5 * dt;
    a.e_f2 = e1;
    a.c_f2 = 0.5 * dt;
    line<kRhsTSZ | kFlagSolve | kFlagLhsJump>(E, 2, a);
    a.e_in = e1;
    a.mib = E.mib[1];
    line<kRhsCopy | kFlagSolve | kFlagLhsJump>(E, 1, a);
    a.mib = E.mib[0];
    line<kRhsCopy | kFlagSolve | kFlagLhsJump>(E, 0, a);
    faces(E, E.u, e1, 0.0);
  } else {
    return 2;
  }
  if (dump) dump_field(dump_dir + "/u1.bin", E.u, nn, E.px);
  std::printf("ok %s mib%d %dx%dx%d crossings %zu %zu %zu regular rows %lld\n", method.c_str(), mib,
              nn[0], nn[1], nn[2], res.cross[0].size(), res.cross[1].size(), res.cross[2].size(),
              E.reg_rows_checked);
  return 0;
}
