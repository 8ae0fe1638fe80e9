// CPU bounds check of the per-step kernels' memory accesses.
//
// Builds the device table for one geometry exactly as an instance does, then
// runs every global-memory and LDS access of one time step of the chosen
// method on the host, using the same per-node / per-lane functions the HIP
// kernels call (device_core.hpp), over buffers of exactly the device sizes.
// Compiled with AddressSanitizer, so any index a kernel would form outside
// its buffer aborts here instead of faulting a GPU.  The tridiagonal solve
// itself touches registers only and is not emulated (rows keep their RHS).
//
//   mts_cpucheck <method adi|lod-ie|lod-cn|ts> <mib 1|2> <nodes per axis>
//                <surface ellipsoid|torus|bumpy> <p0> <p1> [p2 p3]
//                <beta const|sincos> <b0> <b1>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "device_core.hpp"
#include "host_setup.hpp"
#include "host_table.hpp"
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
  constexpr int LPW = 64 / L, T = tile_lines(L);
  (void)E;
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
      for (int l = 0; l < T && j0 + l < nj; l++)
        for (int r = 0; r < n; r++)
          if (seen[(size_t)l * q.lds_len + lds_slot(q, r)] != 1) std::abort();
      std::unique_ptr<double[]> wt(new double[(size_t)T * kWtLen]);
      for (int e = 0; e < T * kWtLen; e++) line_weight_entry<DIR>(p, j0, k, e, wt.get());
      std::vector<double> zt(2 * (size_t)p.g.n[2]);
      for (int e = 0; e < 2 * p.g.n[2]; e++) ztab_entry(p.g, e, zt.data());
      for (int wave = 0; wave < kBlockThreads / 64; wave++)
          for (int lane = 0; lane < 64; lane++) {
            const int kseg = lane & (L - 1);
            const int l = wave * LPW + lane / L;
            const int j = j0 + l;
            const bool valid = j >= 1 && j <= nj - 2;
            const LineMeta lm = line_meta(p.mib, k * nj + j, valid);
            const LaneSeg sg = lane_seg(q, kseg, valid);
            if (sg.m > 0 && (sg.m < 2 || sg.m > MMAX || sg.s < 1 || sg.s + sg.m > n - 1)) std::abort();
            double a[MMAX], b[MMAX], c[MMAX], d[MMAX];
            line_rows<DIR, MMAX, FLAGS>(p, j, k, lm, tA.get() + l * q.lds_len,
                                        tB.get() + l * q.lds_len, wt.get() + l * kWtLen,
                                        DIR == 2 ? zt.data() : nullptr, sg, a, b, c, d);
            double x0 = 0.0, xn = 0.0;
            if (valid) line_ends<DIR>(p, j, k, x0, xn);
            (void)SOLVE;
            if (valid)
              line_write<DIR, MMAX, FLAGS>(p, j, k, lm, sg, d, x0, xn,
                                           (TWO && !HAS_S ? tB.get() : tA.get()) + l * q.lds_len,
                                           DIR == 2 ? zt.data() : nullptr);
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

}  // namespace

int main(int argc, char **argv) {
  if (argc < 9) {
    std::fprintf(stderr, "usage: mts_cpucheck method mib nodes surface p.. beta b0 b1\n");
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
  for (int d = 0; d < 3; d++) {
    E.mib[d].blk_ptr = E.t.blk_ptr[d].data();
    E.mib[d].hdr = E.t.blk_hdr[d].data();
    E.mib[d].val = E.t.blk_val[d].data();
    E.mib[d].cr = E.t.cr_val[d].data();
    E.mib[d].seq = E.t.n_seq[d] > 0 ? E.t.seq[d].data() : nullptr;
    E.mib[d].cls0 = E.t.cls0[d].data();
  }
  LineArgs a;
  std::memset(&a, 0, sizeof(a));
  a.min_lanes = 16;
  if (const char *ev = std::getenv("MTS_LINE_LANES")) a.min_lanes = std::atoi(ev);
  a.g = E.g;
  a.P = E.t.phys;
  a.theta_dt = tdt;
  a.dt = dt;
  a.e_out = a.e_in = a.e_jump = a.e_f1 = a.e_f2 = 0.9;
  a.c_f1 = a.c_f2 = 0.5 * dt;
  if (adi) {
    Rhs1Args r1{E.g, E.t.phys, E.u.data(), E.r.data(), dt, 0.9};
    for (int iz = 0; iz < nn[2]; iz++)
      for (int iy = 0; iy < nn[1]; iy++)
        for (int ix = 0; ix < nn[0]; ix++) rhs1_node(r1, ix, iy, iz);
    SpecArgs sp;
    sp.g = E.g;
    sp.P = E.t.phys;
    for (int d = 0; d < 3; d++) sp.mib[d] = E.mib[d];
    sp.spec = E.t.spec.data();
    sp.nspec = (int)(E.t.spec.size() / kSpecLen);
    sp.u = E.u.data();
    sp.R = E.r.data();
    sp.dt = dt;
    sp.e = 0.9;
    for (int t = 0; t < sp.nspec; t++) spec_node(sp, t);
    a.A = E.r.data();
    a.out = E.r.data();
    a.mib = E.mib[0];
    line<kRhsCopy | kFlagSolve>(E, 0, a);
    a.S = E.u.data();
    a.mib = E.mib[1];
    line<kRhsDouglas | kFlagSolve>(E, 1, a);
    a.mib = E.mib[2];
    a.out = E.u.data();
    line<kRhsDouglas | kFlagSolve>(E, 2, a);
  } else if (method == "lod-ie" || method == "lod-cn") {
    a.A = a.out = E.u.data();
    const bool cn = method == "lod-cn";
    for (int d = 0; d < 3; d++) {
      a.mib = E.mib[d];
      a.in_analytic = d > 0;
      if (cn) {
        if (d < 2) line<kRhsCN | kFlagSolve | kFlagLhsJump>(E, d, a);
        else line<kRhsCN | kFlagSolve | kFlagLhsJump | kFlagEpiF>(E, d, a);
      } else {
        if (d < 2) line<kRhsCopy | kFlagSolve | kFlagLhsJump>(E, d, a);
        else line<kRhsCopy | kFlagSolve | kFlagLhsJump | kFlagEpiF>(E, d, a);
      }
    }
  } else if (method == "ts") {
    a.A = a.out = E.u.data();
    a.mib = E.mib[0];
    line<kRhsCN>(E, 0, a);
    a.in_analytic = 1;
    a.mib = E.mib[1];
    line<kRhsCN>(E, 1, a);
    a.mib = E.mib[2];
    line<kRhsTSZ | kFlagSolve | kFlagLhsJump>(E, 2, a);
    a.mib = E.mib[1];
    line<kRhsCopy | kFlagSolve | kFlagLhsJump>(E, 1, a);
    a.mib = E.mib[0];
    line<kRhsCopy | kFlagSolve | kFlagLhsJump>(E, 0, a);
  } else {
    return 2;
  }
  std::printf("ok %s mib%d %dx%dx%d crossings %zu %zu %zu\n", method.c_str(), mib, nn[0], nn[1], nn[2],
              res.cross[0].size(), res.cross[1].size(), res.cross[2].size());
  return 0;
}
