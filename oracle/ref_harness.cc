// Test-infrastructure harness around the reference solver (Matched-TS).
//
// This file is part of the ORACLE, not of the product.  It is compiled by
// oracle/Makefile together with the reference's own sources where they lie
// under /root/reference/src (nothing is copied) into oracle/_ref/mts_ref, a
// command-line tool that runs the reference's setup and time loop on one
// configuration and dumps what the parity tests compare against:
//   crossings.txt   every interface crossing the reference's Intersections
//                   built (line, ID, left_loc, gamma, coordinates, normal,
//                   P matrix, MIB weights, beta at the crossing)
//   jumps.txt       [u] and u_dir per crossing after the first step's
//                   Refresh_Jump (intersection.cc:2294-2490)
//   u_step1.bin     float64 field after step 1  (x-fastest: (iz*ny+iy)*nx+ix)
//   u_final.bin     float64 field after the last step
//   u_step<j>.bin   with --dump-steps K: the field after step j, j = 1..K
//   summary.txt     step count, final time, Lmax/L2/argmax (same loop as
//                   Douglas_ADI::Error, douglas_adi.cc:1723-1770), timings
//
// The reference has no plugins for the manufactured solution, diffusion
// coefficients and bumpy sphere its benchmark configurations use, so this
// harness defines them behind the reference's own abstract interfaces
// (Equation equation.h:5-37, Beta beta.h:5-15, Surface_Cartesian
// surface_cartesian.h:6-26).  The product implements the same formulas in
// the same floating-point order.
//
// The time loop follows the reference drivers helper.h:199-238 (ADI),
// helper.h:280-319 (LOD) and helper.h:361-401 (TS): loop = (int)(T/dt),
// t += dt, a fresh Equation per step.

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "constant.h"
#include "diffusion/beta.h"
#include "equation/equation.h"
#include "mesh/mesh.h"
#include "spatial/intersection.h"
#include "surface/surface_cartesian.h"
#include "surface/surface_ellipsoid.h"
#include "surface/surface_torus.h"
#include "temporal/douglas_adi.h"
#include "temporal/lod.h"
#include "temporal/trapezoidal.h"

using namespace std;

// ---------------------------------------------------------------- beta kinds
// beta- = b_in, beta+ = b_out (constants).
class HBetaConst : public Beta {
  double bin_, bout_;

 public:
  HBetaConst(double bin, double bout) : bin_(bin), bout_(bout) {}
  double Inside(Doub_I, Doub_I, Doub_I) const override { return bin_; }
  double Outside(Doub_I, Doub_I, Doub_I) const override { return bout_; }
  double Inside_Dx(Doub_I, Doub_I, Doub_I) const override { return 0; }
  double Outside_Dx(Doub_I, Doub_I, Doub_I) const override { return 0; }
  double Inside_Dy(Doub_I, Doub_I, Doub_I) const override { return 0; }
  double Outside_Dy(Doub_I, Doub_I, Doub_I) const override { return 0; }
  double Inside_Dz(Doub_I, Doub_I, Doub_I) const override { return 0; }
  double Outside_Dz(Doub_I, Doub_I, Doub_I) const override { return 0; }
};

// beta- = c (1 + 0.5 sin x cos y), beta+ = b_out.
class HBetaSinCos : public Beta {
  double c_, bout_;

 public:
  HBetaSinCos(double c, double bout) : c_(c), bout_(bout) {}
  double Inside(Doub_I x, Doub_I y, Doub_I) const override {
    return c_ * (1.0 + 0.5 * sin(x) * cos(y));
  }
  double Outside(Doub_I, Doub_I, Doub_I) const override { return bout_; }
  double Inside_Dx(Doub_I x, Doub_I y, Doub_I) const override {
    return 0.5 * c_ * cos(x) * cos(y);
  }
  double Outside_Dx(Doub_I, Doub_I, Doub_I) const override { return 0; }
  double Inside_Dy(Doub_I x, Doub_I y, Doub_I) const override {
    return -0.5 * c_ * sin(x) * sin(y);
  }
  double Outside_Dy(Doub_I, Doub_I, Doub_I) const override { return 0; }
  double Inside_Dz(Doub_I, Doub_I, Doub_I) const override { return 0; }
  double Outside_Dz(Doub_I, Doub_I, Doub_I) const override { return 0; }
};

// ------------------------------------------------ manufactured solution
// u- = e^{-t} sin x sin y sin z, u+ = e^{-t} cos x cos y cos z,
// f = u_t - div(beta grad u) = (3 beta - 1) u - grad(beta).grad(u).
class HEqExp : public Equation {
  Beta &beta;
  double e_;  // exp(-t)

  double Inner_dux(Doub_I x, Doub_I y, Doub_I z) const override {
    return e_ * cos(x) * sin(y) * sin(z);
  }
  double Inner_duy(Doub_I x, Doub_I y, Doub_I z) const override {
    return e_ * sin(x) * cos(y) * sin(z);
  }
  double Inner_duz(Doub_I x, Doub_I y, Doub_I z) const override {
    return e_ * sin(x) * sin(y) * cos(z);
  }
  double Outer_dux(Doub_I x, Doub_I y, Doub_I z) const override {
    return -e_ * sin(x) * cos(y) * cos(z);
  }
  double Outer_duy(Doub_I x, Doub_I y, Doub_I z) const override {
    return -e_ * cos(x) * sin(y) * cos(z);
  }
  double Outer_duz(Doub_I x, Doub_I y, Doub_I z) const override {
    return -e_ * cos(x) * cos(y) * sin(z);
  }

 public:
  HEqExp(double ti, Beta &b) : beta(b) {
    t = ti;
    e_ = exp(-ti);
  }
  double Outer_u(Doub_I x, Doub_I y, Doub_I z) const override {
    return e_ * cos(x) * cos(y) * cos(z);
  }
  double Inner_u(Doub_I x, Doub_I y, Doub_I z) const override {
    return e_ * sin(x) * sin(y) * sin(z);
  }
  double Outer_f(Doub_I x, Doub_I y, Doub_I z) const override {
    return (3.0 * beta.Outside(x, y, z) - 1.0) * Outer_u(x, y, z) -
           beta.Outside_Dx(x, y, z) * Outer_dux(x, y, z) -
           beta.Outside_Dy(x, y, z) * Outer_duy(x, y, z) -
           beta.Outside_Dz(x, y, z) * Outer_duz(x, y, z);
  }
  double Inner_f(Doub_I x, Doub_I y, Doub_I z) const override {
    return (3.0 * beta.Inside(x, y, z) - 1.0) * Inner_u(x, y, z) -
           beta.Inside_Dx(x, y, z) * Inner_dux(x, y, z) -
           beta.Inside_Dy(x, y, z) * Inner_duy(x, y, z) -
           beta.Inside_Dz(x, y, z) * Inner_duz(x, y, z);
  }
  double Jump_betau_x(Doub_I x, Doub_I y, Doub_I z) const override {
    return beta.Outside(x, y, z) * Outer_dux(x, y, z) -
           beta.Inside(x, y, z) * Inner_dux(x, y, z);
  }
  double Jump_betau_y(Doub_I x, Doub_I y, Doub_I z) const override {
    return beta.Outside(x, y, z) * Outer_duy(x, y, z) -
           beta.Inside(x, y, z) * Inner_duy(x, y, z);
  }
  double Jump_betau_z(Doub_I x, Doub_I y, Doub_I z) const override {
    return beta.Outside(x, y, z) * Outer_duz(x, y, z) -
           beta.Inside(x, y, z) * Inner_duz(x, y, z);
  }
};

// ------------------------------------------------------------ bumpy sphere
// phi = |x| - (R0 + A sin(m theta) cos(k psi)), theta = atan2(rho, z),
// psi = atan2(y, x), rho = sqrt(x^2 + y^2).  On the z axis (rho == 0) the
// psi-dependent terms have no limit; there the gradient is taken radial.
class HSurfBumpy : public Surface_Cartesian {
  double R0_, A_, m_, k_;

  void grad(double x, double y, double z, double g[3]) const {
    double rho2 = x * x + y * y;
    double r2 = rho2 + z * z;
    double r = sqrt(r2);
    double rho = sqrt(rho2);
    if (rho == 0.0) {
      g[0] = 0.0;
      g[1] = 0.0;
      g[2] = z / r;
      return;
    }
    double th = atan2(rho, z);
    double ps = atan2(y, x);
    double dRdth = A_ * m_ * cos(m_ * th) * cos(k_ * ps);
    double dRdps = -A_ * k_ * sin(m_ * th) * sin(k_ * ps);
    double dthdx = x * z / (r2 * rho);
    double dthdy = y * z / (r2 * rho);
    double dthdz = -rho / r2;
    double dpsdx = -y / rho2;
    double dpsdy = x / rho2;
    g[0] = x / r - (dRdth * dthdx + dRdps * dpsdx);
    g[1] = y / r - (dRdth * dthdy + dRdps * dpsdy);
    g[2] = z / r - dRdth * dthdz;
  }
  double func(const Doub_I x, const Doub_I y, const Doub_I z) const override {
    double rho = sqrt(x * x + y * y);
    double r = sqrt(x * x + y * y + z * z);
    double th = atan2(rho, z);
    double ps = atan2(y, x);
    return r - (R0_ + A_ * sin(m_ * th) * cos(k_ * ps));
  }
  double dfx(const Doub_I x, const Doub_I y, const Doub_I z) const override {
    double g[3];
    grad(x, y, z, g);
    return g[0];
  }
  double dfy(const Doub_I x, const Doub_I y, const Doub_I z) const override {
    double g[3];
    grad(x, y, z, g);
    return g[1];
  }
  double dfz(const Doub_I x, const Doub_I y, const Doub_I z) const override {
    double g[3];
    grad(x, y, z, g);
    return g[2];
  }

 public:
  HSurfBumpy(double R0, double A, double m, double k)
      : R0_(R0), A_(A), m_(m), k_(k) {}
  double Set_Node(Doub_I x, Doub_I y, Doub_I z) const override {
    return func(x, y, z) + TOL_SETUP;
  }
  void display() override {}
  VecDoub normal(Doub_I x, Doub_I y, Doub_I z) override {
    double g[3];
    grad(x, y, z, g);
    double tot = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    return VecDoub{g[0] / tot, g[1] / tot, g[2] / tot};
  }
  double check(Doub_I x, Doub_I y, Doub_I z) override { return func(x, y, z); }
};

// ------------------------------------------------------------------ options
struct Opts {
  string method = "adi";
  int mib = 1;
  int n = 32;  // intervals per axis
  double lo = -1.0, hi = 1.0;
  double dt = 0.0625, T = 1.0;
  string surface = "ellipsoid:1,1,1,0.5";
  string beta = "const:10,1";
  string out = ".";
  int max_steps = -1;  // cap on steps (timing samples); -1 = all
  int dump = 1;
  int dump_steps = 0;  // fields after steps 1..dump_steps (u_step<j>.bin)
};

static vector<double> parse_list(const string &s) {
  vector<double> v;
  size_t p = 0;
  while (p < s.size()) {
    size_t q = s.find(',', p);
    if (q == string::npos) q = s.size();
    v.push_back(strtod(s.substr(p, q - p).c_str(), nullptr));
    p = q + 1;
  }
  return v;
}

static void usage() {
  fprintf(stderr,
          "mts_ref --method adi|lod-ie|lod-cn|ts --mib 1|2 --n N --lo L --hi H\n"
          "        --dt DT --T T --surface ellipsoid:a,b,c,r|torus:R,r|bumpy:R0,A,m,k\n"
          "        --beta const:bin,bout|sincos:c,bout --out DIR [--max-steps K] [--no-dump]\n"
          "        [--dump-steps K]\n");
  exit(2);
}

static void dump_field(const string &path, const CubicDoub &uh, int nx, int ny,
                       int nz) {
  vector<double> buf((size_t)nx * ny * nz);
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++)
        buf[((size_t)iz * ny + iy) * nx + ix] = uh[ix][iy][iz];
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) {
    perror(path.c_str());
    exit(1);
  }
  fwrite(buf.data(), sizeof(double), buf.size(), f);
  fclose(f);
}

static void dump_crossing(FILE *f, const Intersection_Data &d, int j, int k) {
  fprintf(f, "%c %d %d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g",
          d.dir, j, k, d.ID, d.left_loc, d.gamma, d.coord.x_value,
          d.coord.y_value, d.coord.z_value, d.local.normal[0],
          d.local.normal[1], d.local.normal[2], d.diffcoef.inside,
          d.diffcoef.outside);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) fprintf(f, " %.17g", d.p[a][b]);
  size_t nw = d.wei.weil[0].size();
  fprintf(f, " %zu", nw);
  for (size_t i = 0; i < nw; i++) fprintf(f, " %.17g", d.wei.weil[0][i]);
  for (size_t i = 0; i < nw; i++) fprintf(f, " %.17g", d.wei.weir[0][i]);
  fprintf(f, "\n");
}

template <class F>
static void for_each_crossing(Intersections &inter, int nx, int ny, int nz, F fn) {
  for (int ix = 1; ix < nx - 1; ix++)
    for (int iy = 1; iy < ny - 1; iy++)
      for (auto &d : inter.ifpz[ix][iy]) fn(d, ix, iy);
  for (int ix = 1; ix < nx - 1; ix++)
    for (int iz = 1; iz < nz - 1; iz++)
      for (auto &d : inter.ifpy[ix][iz]) fn(d, ix, iz);
  for (int iy = 1; iy < ny - 1; iy++)
    for (int iz = 1; iz < nz - 1; iz++)
      for (auto &d : inter.ifpx[iy][iz]) fn(d, iy, iz);
}

int main(int argc, char **argv) {
  Opts o;
  for (int i = 1; i < argc; i++) {
    string a = argv[i];
    auto need = [&]() -> string {
      if (i + 1 >= argc) usage();
      return string(argv[++i]);
    };
    if (a == "--method") o.method = need();
    else if (a == "--mib") o.mib = atoi(need().c_str());
    else if (a == "--n") o.n = atoi(need().c_str());
    else if (a == "--lo") o.lo = strtod(need().c_str(), nullptr);
    else if (a == "--hi") o.hi = strtod(need().c_str(), nullptr);
    else if (a == "--dt") o.dt = strtod(need().c_str(), nullptr);
    else if (a == "--T") o.T = strtod(need().c_str(), nullptr);
    else if (a == "--surface") o.surface = need();
    else if (a == "--beta") o.beta = need();
    else if (a == "--out") o.out = need();
    else if (a == "--max-steps") o.max_steps = atoi(need().c_str());
    else if (a == "--no-dump") o.dump = 0;
    else if (a == "--dump-steps") o.dump_steps = atoi(need().c_str());
    else usage();
  }

  // plugins
  unique_ptr<Beta> beta;
  {
    size_t c = o.beta.find(':');
    string kind = o.beta.substr(0, c);
    vector<double> p = parse_list(o.beta.substr(c + 1));
    if (kind == "const") beta.reset(new HBetaConst(p[0], p[1]));
    else if (kind == "sincos") beta.reset(new HBetaSinCos(p[0], p[1]));
    else usage();
  }
  unique_ptr<Surface_Cartesian> surf;
  {
    size_t c = o.surface.find(':');
    string kind = o.surface.substr(0, c);
    vector<double> p = parse_list(o.surface.substr(c + 1));
    if (kind == "ellipsoid") surf.reset(new Surface_Ellipsoid(p));
    else if (kind == "torus") surf.reset(new Surface_Torus(p));
    else if (kind == "bumpy") surf.reset(new HSurfBumpy(p[0], p[1], p[2], p[3]));
    else usage();
  }

  VecDoub domain{o.lo, o.hi, o.lo, o.hi, o.lo, o.hi};
  VecInt size{o.n + 1, o.n + 1, o.n + 1};
  VecDoub time{0.0, o.T, o.dt};
  int accuracy = 2;

  ofstream log(o.out + "/reference_log.txt", ios::out);
  auto t0 = chrono::steady_clock::now();
  Mesh mesh(domain, size, *surf);
  Intersections inter(*surf, mesh, *beta, accuracy, o.mib, log);
  auto t1 = chrono::steady_clock::now();
  int nx = mesh.nx, ny = mesh.ny, nz = mesh.nz;

  if (o.dump) {
    FILE *f = fopen((o.out + "/crossings.txt").c_str(), "w");
    for_each_crossing(inter, nx, ny, nz, [&](Intersection_Data &d, int j, int k) {
      dump_crossing(f, d, j, k);
    });
    fclose(f);
  }

  CubicDoub uh;
  double dt = time[2], t_now = time[0], t_finish = time[1];
  int loop = (int)(t_finish / dt) / NPRINT;
  if (o.max_steps >= 0 && loop > o.max_steps) loop = o.max_steps;

  unique_ptr<Douglas_ADI> adi;
  unique_ptr<LOD> lod;
  unique_ptr<TS> ts;
  char lod_method = 'C';
  {
    HEqExp eq0(time[0], *beta);
    if (o.method == "adi") {
      adi.reset(new Douglas_ADI(inter, mesh, *beta, time, accuracy));
      adi->Initialization(eq0, uh);
    } else if (o.method == "lod-ie" || o.method == "lod-cn") {
      lod_method = (o.method == "lod-ie") ? 'I' : 'C';
      lod.reset(new LOD(inter, mesh, *beta, time));
      lod->Initialization(eq0, uh);
    } else if (o.method == "ts") {
      ts.reset(new TS(inter, mesh, *beta, time));
      ts->Initialization(eq0, uh);
    } else {
      usage();
    }
  }

  auto t2 = chrono::steady_clock::now();
  double first_step_s = 0.0;
  for (int j = 0; j < loop; j++) {
    auto ts0 = chrono::steady_clock::now();
    t_now += dt;
    if (adi) {
      HEqExp eq(t_now, *beta);
      adi->Solve_2nd(eq, inter, uh, *beta);
    } else if (lod) {
      HEqExp eq(t_now, *beta);
      lod->Solve_2nd(eq, inter, uh, *beta, lod_method);
    } else {
      HEqExp eqn(t_now - dt, *beta), eqh(t_now - 0.5 * dt, *beta), eq1(t_now, *beta);
      ts->Solve_2nd(eqn, eqh, eq1, inter, uh, *beta);
    }
    if (o.dump && j < o.dump_steps)
      dump_field(o.out + "/u_step" + to_string(j + 1) + ".bin", uh, nx, ny, nz);
    if (j == 0) {
      first_step_s = chrono::duration<double>(chrono::steady_clock::now() - ts0).count();
      if (o.dump) {
        dump_field(o.out + "/u_step1.bin", uh, nx, ny, nz);
        FILE *f = fopen((o.out + "/jumps.txt").c_str(), "w");
        for_each_crossing(inter, nx, ny, nz, [&](Intersection_Data &d, int, int) {
          fprintf(f, "%.17g %.17g\n", d.jump.u, d.jump.u_dir);
        });
        fclose(f);
      }
    }
  }
  auto t3 = chrono::steady_clock::now();
  if (o.dump) dump_field(o.out + "/u_final.bin", uh, nx, ny, nz);

  // error norms, same loop and formulas as Douglas_ADI::Error
  HEqExp eqT(t_now, *beta);
  double lmax = 0.0, l2 = 0.0;
  int xm = 0, ym = 0, zm = 0;
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++) {
        int indx = iz * nx * ny + iy * nx + ix;
        double ex = (mesh.mesh_value[indx] == 1)
                        ? eqT.Outer_u(mesh.xi[ix], mesh.yi[iy], mesh.zi[iz])
                        : eqT.Inner_u(mesh.xi[ix], mesh.yi[iy], mesh.zi[iz]);
        double e = fabs(uh[ix][iy][iz] - ex);
        if (e > lmax) {
          lmax = e;
          xm = ix;
          ym = iy;
          zm = iz;
        }
        l2 += e * e;
      }
  l2 = sqrt(l2 / (1.0 * nx * ny * nz));

  long ncross[3] = {0, 0, 0};
  for_each_crossing(inter, nx, ny, nz, [&](Intersection_Data &d, int, int) {
    ncross[d.dir - 'x']++;
  });

  // end-of-run diagnostics at the final time, as the reference's driver
  // reports them (helper.h:222-236): fictitious-point errors (Refresh_Fp,
  // intersection.cc:1841) and jump errors (Refresh_Jump), per crossing, plus
  // the reference's own Error_Fp / Error_Jump report text
  if (o.dump) {
    inter.Refresh_Fp(eqT);
    inter.Refresh_Jump(eqT, uh, *beta);
    FILE *fd = fopen((o.out + "/diag.txt").c_str(), "w");
    for_each_crossing(inter, nx, ny, nz, [&](Intersection_Data &d, int, int) {
      fprintf(fd, "%.17g %.17g %.17g\n", d.err.errl.empty() ? 0.0 : d.err.errl[0],
              d.err.errr.empty() ? 0.0 : d.err.errr[0], d.jump.err);
    });
    fclose(fd);
    ofstream rep(o.out + "/diag_report.txt");
    inter.Error_Fp(rep);
    inter.Error_Jump(rep);
  }

  FILE *f = fopen((o.out + "/summary.txt").c_str(), "w");
  fprintf(f, "nx %d\nny %d\nnz %d\nsteps %d\nt_final %.17g\n", nx, ny, nz, loop, t_now);
  fprintf(f, "lmax %.17g\nl2 %.17g\nargmax %d %d %d\n", lmax, l2, xm, ym, zm);
  fprintf(f, "ncross_x %ld\nncross_y %ld\nncross_z %ld\n", ncross[0], ncross[1], ncross[2]);
  fprintf(f, "setup_s %.6f\nstep_loop_s %.6f\nfirst_step_s %.6f\n",
          chrono::duration<double>(t1 - t0).count(),
          chrono::duration<double>(t3 - t2).count(), first_step_s);
  fclose(f);
  return 0;
}
