/*
 * C++ drop-in classes over the C ABI (include/mts.h), with the reference's
 * class surface: Mesh, Intersections, Douglas_ADI, LOD, TS and LU, with the
 * reference's constructor and method signatures (paths relative to the
 * reference's src/):
 *
 *   Mesh(const VecDoub_I, const VecInt_I, Surface_Cartesian &)   mesh/mesh.h:20
 *   Intersections(Surface_Cartesian &, Mesh &, Beta &, Int_I accuracy,
 *                 Int_I mib_method, ofstream &)                  spatial/intersection.h:88
 *   Douglas_ADI(Intersections &, Mesh &, Beta &, VecDoub_I time, Int_I)
 *     Solve_2nd(Equation &, Intersections &, CubicDoub_I &, Beta &)
 *     Initialization(Equation &, CubicDoub_I &)
 *     Error(Equation &, CubicDoub_I &, ofstream &)              temporal/douglas_adi.h:47-52
 *   LOD(Intersections &, Mesh &, Beta &, VecDoub_I)
 *     Solve_2nd(Equation &, Intersections &, CubicDoub_I &, Beta &, char &)
 *                                                                temporal/lod.h:51-56
 *   TS(Intersections &, Mesh &, Beta &, VecDoub_I)
 *     Solve_2nd(Equation &, Equation &, Equation &, Intersections &,
 *               CubicDoub_I &, Beta &)                           temporal/trapezoidal.h:47-53
 *   LU(MatrixDoub_I &): solve, inverse, det, mprove              decomposition/lu.h:7-21
 *
 * The per-step path runs on the GPU through the C ABI; these classes hold
 * handles and copy fields, nothing more.  Plugins: the surfaces are the
 * library's built-in level sets (ellipsoid, torus, bumpy sphere) or any
 * Surface_Cartesian subclass that defines func and its gradient (called on
 * the host during setup); Beta is constant or the sin-cos variable form; the
 * Equation is the manufactured solution u- = e^-t sin x sin y sin z,
 * u+ = e^-t cos x cos y cos z that the kernels evaluate on the device.
 *
 * Errors: where the reference calls exit() or throws, these classes throw
 * mts::Error (the library's message and MTS_E_* status).  With
 * mts::set_exit_on_error(true), or MTS_EXIT_ON_ERROR=1 in the environment,
 * they print "[ERROR]: <message>" to stderr and exit(EXIT_FAILURE) instead.
 *
 * Host copies: Solve_2nd updates uh in place, as the reference does (upload
 * before and download after every step).  A caller that only reads uh at the
 * end can call set_host_sync(false) on the solver: the field then stays on
 * the device between steps and uh is refreshed by Initialization, Error and
 * sync_host(uh).
 */
#ifndef MTS_HPP
#define MTS_HPP

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "mts.h"

namespace mts {

// the reference's container typedefs (constant.h:18-27)
typedef int Int_I, Int_O, Int_IO;
typedef double Doub_I, Doub_O, Doub_IO;
typedef std::vector<int> VecInt, VecInt_I, VecInt_O, VecInt_IO;
typedef std::vector<double> VecDoub, VecDoub_I, VecDoub_O, VecDoub_IO;
typedef std::vector<std::vector<double>> MatrixDoub, MatrixDoub_I, MatrixDoub_O, MatrixDoub_IO;
typedef std::vector<std::vector<std::vector<double>>> CubicDoub, CubicDoub_I, CubicDoub_O,
    CubicDoub_IO;

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline bool &exit_on_error_flag() {
  static bool on = [] {
    const char *e = std::getenv("MTS_EXIT_ON_ERROR");
    return e && e[0] == '1';
  }();
  return on;
}
inline void set_exit_on_error(bool on) { exit_on_error_flag() = on; }

inline void fail(int code, const std::string &msg) {
  if (exit_on_error_flag()) {
    std::cerr << "[ERROR]: " << msg << std::endl;
    std::exit(EXIT_FAILURE);
  }
  throw Error(code, msg);
}
inline void check(int rc) {
  if (rc != MTS_OK) fail(rc, mts_last_error());
}

// ------------------------------------------------------------- surfaces
// Level set phi, > 0 outside, < 0 inside (surface/surface_cartesian.h:6-26).
class Surface_Cartesian {
 public:
  virtual ~Surface_Cartesian() = default;
  virtual double func(double x, double y, double z) = 0;
  virtual double dfx(double x, double y, double z) = 0;
  virtual double dfy(double x, double y, double z) = 0;
  virtual double dfz(double x, double y, double z) = 0;
  // The library's view of the surface: a built-in kind, or func and its
  // gradient called back on the host.
  virtual mts_surface_desc desc() {
    mts_surface_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = MTS_SURF_CALLBACK;
    d.phi = [](void *u, double x, double y, double z) {
      return static_cast<Surface_Cartesian *>(u)->func(x, y, z);
    };
    d.grad = [](void *u, double x, double y, double z, double g[3]) {
      Surface_Cartesian *s = static_cast<Surface_Cartesian *>(u);
      g[0] = s->dfx(x, y, z);
      g[1] = s->dfy(x, y, z);
      g[2] = s->dfz(x, y, z);
    };
    d.user = this;
    return d;
  }
};

// x^2/a^2 + y^2/b^2 + z^2/c^2 - r^2 (surface/surface_ellipsoid.cc; args {a, b, c, r})
class Surface_Ellipsoid : public Surface_Cartesian {
 public:
  explicit Surface_Ellipsoid(VecDoub_I arg) : a_(arg.at(0)), b_(arg.at(1)), c_(arg.at(2)), r_(arg.at(3)) {}
  double func(double x, double y, double z) override {
    return x * x / (a_ * a_) + y * y / (b_ * b_) + z * z / (c_ * c_) - r_ * r_;
  }
  double dfx(double x, double, double) override { return 2 * x / (a_ * a_); }
  double dfy(double, double y, double) override { return 2 * y / (b_ * b_); }
  double dfz(double, double, double z) override { return 2 * z / (c_ * c_); }
  mts_surface_desc desc() override {
    mts_surface_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = MTS_SURF_ELLIPSOID;
    d.p[0] = a_, d.p[1] = b_, d.p[2] = c_, d.p[3] = r_;
    return d;
  }

 private:
  double a_, b_, c_, r_;
};

// (R - sqrt(x^2 + y^2))^2 + z^2 - r^2 (surface/surface_torus.cc; args {R, r})
class Surface_Torus : public Surface_Cartesian {
 public:
  explicit Surface_Torus(VecDoub_I arg) : R_(arg.at(0)), r_(arg.at(1)) {}
  double func(double x, double y, double z) override {
    const double q = R_ - std::sqrt(x * x + y * y);
    return q * q + z * z - r_ * r_;
  }
  double dfx(double x, double y, double) override {
    const double rho = std::sqrt(x * x + y * y);
    return -2 * (R_ - rho) * x / rho;
  }
  double dfy(double x, double y, double) override {
    const double rho = std::sqrt(x * x + y * y);
    return -2 * (R_ - rho) * y / rho;
  }
  double dfz(double, double, double z) override { return 2 * z; }
  mts_surface_desc desc() override {
    mts_surface_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = MTS_SURF_TORUS;
    d.p[0] = R_, d.p[1] = r_;
    return d;
  }

 private:
  double R_, r_;
};

// |x| - (R0 + A sin(m theta) cos(k psi)) (args {R0, A, m, k}); the library
// evaluates it and its gradient itself.
class Surface_Bumpy : public Surface_Cartesian {
 public:
  explicit Surface_Bumpy(VecDoub_I arg) : p_(arg) { p_.resize(4); }
  double func(double x, double y, double z) override {
    const double rho = std::sqrt(x * x + y * y), r = std::sqrt(rho * rho + z * z);
    return r - (p_[0] + p_[1] * std::sin(p_[2] * std::atan2(rho, z)) *
                            std::cos(p_[3] * std::atan2(y, x)));
  }
  double dfx(double, double, double) override { return 0.0; }  // evaluated by the library
  double dfy(double, double, double) override { return 0.0; }
  double dfz(double, double, double) override { return 0.0; }
  mts_surface_desc desc() override {
    mts_surface_desc d;
    std::memset(&d, 0, sizeof(d));
    d.kind = MTS_SURF_BUMPY;
    for (int i = 0; i < 4; i++) d.p[i] = p_[i];
    return d;
  }

 private:
  VecDoub p_;
};

// ------------------------------------------------------------ diffusion
// beta- inside, beta+ outside (diffusion/beta.h:5-15); the forms the kernels evaluate.
class Beta {
 public:
  Beta(int kind, double minus, double plus) {
    std::memset(&d_, 0, sizeof(d_));
    d_.beta_kind = kind;
    d_.bp[0] = minus;
    d_.bp[1] = plus;
  }
  virtual ~Beta() = default;
  double Inside(double x, double y, double) const {
    return d_.beta_kind == MTS_BETA_SINCOS ? d_.bp[0] * (1.0 + 0.5 * std::sin(x) * std::cos(y))
                                           : d_.bp[0];
  }
  double Outside(double, double, double) const { return d_.bp[1]; }
  const mts_physics_desc &desc() const { return d_; }

 private:
  mts_physics_desc d_;
};
// beta- = minus, beta+ = plus
class Beta_Const : public Beta {
 public:
  Beta_Const(double minus, double plus) : Beta(MTS_BETA_CONST, minus, plus) {}
};
// beta- = c (1 + 0.5 sin x cos y), beta+ = plus
class Beta_SinCos : public Beta {
 public:
  Beta_SinCos(double c, double plus) : Beta(MTS_BETA_SINCOS, c, plus) {}
};

// ------------------------------------------------------------- equation
// The equation at time t (equation/equation.h:5-37): u- = e^-t sin x sin y
// sin z, u+ = e^-t cos x cos y cos z, f = u_t - div(beta grad u) per side.
class Equation {
 public:
  Equation(double t, Beta &beta) : t_(t), beta_(&beta) {}
  double t() const { return t_; }
  Beta &beta() const { return *beta_; }
  double Inner_u(double x, double y, double z) const {
    return std::exp(-t_) * std::sin(x) * std::sin(y) * std::sin(z);
  }
  double Outer_u(double x, double y, double z) const {
    return std::exp(-t_) * std::cos(x) * std::cos(y) * std::cos(z);
  }

 private:
  double t_;
  Beta *beta_;
};

// --------------------------------------------------------------- mesh
// Grid and node classification (mesh/mesh.h:10-25).  xi, yi, zi and
// mesh_value point into storage the object owns (the reference leaks it).
class Mesh {
 public:
  double xl, xr, yl, yr, zl, zr, dx, dy, dz;
  int nx, ny, nz;
  double *xi, *yi, *zi, *mesh_value;

  Mesh(const VecDoub_I domain, const VecInt_I size, Surface_Cartesian &ex) : surface_(&ex) {
    if (domain.size() < 6 || size.size() < 3) fail(MTS_E_ARG, "Mesh: domain needs 6 bounds and size 3 counts");
    xl = domain[0], xr = domain[1], yl = domain[2], yr = domain[3], zl = domain[4], zr = domain[5];
    nx = size[0], ny = size[1], nz = size[2];
    dx = (xr - xl) / (nx - 1.0);
    dy = (yr - yl) / (ny - 1.0);
    dz = (zr - zl) / (nz - 1.0);
    x_.resize(nx), y_.resize(ny), z_.resize(nz);
    for (int i = 0; i < nx; i++) x_[i] = xl + i * dx;
    for (int i = 0; i < ny; i++) y_[i] = yl + i * dy;
    for (int i = 0; i < nz; i++) z_[i] = zl + i * dz;
    const mts_grid_desc g = grid();
    const mts_surface_desc s = ex.desc();
    std::vector<int8_t> ind((size_t)nx * ny * nz);
    check(mts_mesh_classify(&g, &s, ind.data(), ind.size()));
    v_.assign(ind.begin(), ind.end());
    xi = x_.data(), yi = y_.data(), zi = z_.data(), mesh_value = v_.data();
  }
  Mesh(const Mesh &) = delete;
  Mesh &operator=(const Mesh &) = delete;

  mts_grid_desc grid() const {
    mts_grid_desc g;
    g.lo[0] = xl, g.lo[1] = yl, g.lo[2] = zl;
    g.hi[0] = xr, g.hi[1] = yr, g.hi[2] = zr;
    g.n[0] = nx, g.n[1] = ny, g.n[2] = nz;
    return g;
  }
  Surface_Cartesian &surface() const { return *surface_; }
  void display() const {
    std::cout << "Mesh: [" << xl << "," << xr << "]x[" << yl << "," << yr << "]x[" << zl << ","
              << zr << "], " << nx << "x" << ny << "x" << nz << " nodes" << std::endl;
  }

 private:
  Surface_Cartesian *surface_;
  VecDoub x_, y_, z_, v_;
};

// -------------------------------------------------------- intersections
// Crossings and MIB weights (spatial/intersection.h:84-104), built by the
// library once in FP64.  ifpx[iy][iz], ifpy[ix][iz], ifpz[ix][iy] list the
// crossings of each line in order.  The jumps are analytic and evaluated by
// the kernels at every step; Refresh_Jump only sets the diagnostics' time.
class Intersections {
 public:
  typedef std::vector<std::vector<std::vector<mts_crossing>>> Cubic_Intersection_Data;
  Cubic_Intersection_Data ifpx, ifpy, ifpz;

  Intersections(Surface_Cartesian &ex, Mesh &mesh, Beta &beta, Int_I accuracy, Int_I mib_method,
                std::ofstream &) {
    const mts_grid_desc g = mesh.grid();
    const mts_surface_desc s = ex.desc();
    phys_ = beta.desc();
    mts_setup *st = nullptr;
    check(mts_setup_create(&g, &s, &phys_, mib_method, accuracy, &st));
    setup_.reset(st);
    mts_setup_info info;
    check(mts_setup_get_info(st, &info));
    Cubic_Intersection_Data *out[3] = {&ifpx, &ifpy, &ifpz};
    const int dims[3][2] = {{g.n[1], g.n[2]}, {g.n[0], g.n[2]}, {g.n[0], g.n[1]}};
    for (int d = 0; d < 3; d++) {
      out[d]->assign(dims[d][0], std::vector<std::vector<mts_crossing>>(dims[d][1]));
      std::vector<mts_crossing> c((size_t)info.n_cross[d]);
      if (!c.empty()) check(mts_setup_crossings(st, d, c.data(), c.size()));
      for (const mts_crossing &x : c) (*out[d])[x.j][x.k].push_back(x);
    }
  }
  // The end-of-run diagnostics at the equation's time (intersection.h:95-98):
  // Refresh_* take the time, Error_* print the reference's report blocks.
  void Refresh_Fp(Equation &eq) { t_fp_ = eq.t(); }
  void Refresh_Jump(Equation &eq, CubicDoub &, Beta &) { t_jump_ = eq.t(); }
  void Error_Fp(std::ofstream &out) {
    double fp[2], jp[2];
    check(mts_setup_diagnostics(setup_.get(), t_fp_, fp, jp));
    report(out, "------------------------ Error of FPs ------------------------", fp);
  }
  void Error_Jump(std::ofstream &out) {
    double fp[2], jp[2];
    check(mts_setup_diagnostics(setup_.get(), t_jump_, fp, jp));
    report(out, "------------------------ Error of Jumps ------------------------", jp);
  }
  const mts_setup *handle() const { return setup_.get(); }
  const mts_physics_desc &physics() const { return phys_; }

 private:
  struct Del {
    void operator()(mts_setup *s) const { mts_setup_destroy(s); }
  };
  static void report(std::ofstream &out, const char *title, const double v[2]) {
    out << title << std::endl;
    out << std::setprecision(3) << std::scientific;
    out << "Lmax = " << v[0] << std::endl;
    out << "L2   = " << v[1] << std::endl;
    out << std::fixed << std::endl;
  }
  std::unique_ptr<mts_setup, Del> setup_;
  mts_physics_desc phys_;
  double t_fp_ = 0.0, t_jump_ = 0.0;
};

// ------------------------------------------------------ temporal solvers
namespace detail {
// One context per process on the device MTS_DEVICE (default 0), kept until exit.
inline mts_ctx *context() {
  static mts_ctx *ctx = [] {
    mts_ctx *c = nullptr;
    const char *e = std::getenv("MTS_DEVICE");
    check(mts_ctx_create(e ? std::atoi(e) : 0, &c));
    return c;
  }();
  return ctx;
}

class Solver {
 public:
  Solver(Intersections &inter, Mesh &mesh, VecDoub_I time)
      : inter_(&inter), nx_(mesh.nx), ny_(mesh.ny), nz_(mesh.nz) {
    if (time.size() < 3) fail(MTS_E_ARG, "time needs {start, finish, step}");
    dt_ = time[2];
  }
  virtual ~Solver() {
    if (inst_) mts_instance_destroy(inst_);
  }
  Solver(const Solver &) = delete;
  Solver &operator=(const Solver &) = delete;

  // false: keep the field on the device between steps (see the file comment)
  void set_host_sync(bool on) { host_sync_ = on; }
  void sync_host(CubicDoub &uh) { download(uh); }
  mts_instance *handle() const { return inst_; }

  void Initialization(Equation &eq, CubicDoub &uh) {
    ensure(method_or_default());
    check(mts_initialize(inst_, eq.t()));
    download(uh);
  }
  // Error (douglas_adi.cc:1723-1770): L-inf with its first argmax and L2 over all nodes.
  void Error(Equation &eq, CubicDoub &uh, std::ofstream &out) {
    double lmax = 0.0, l2 = 0.0;
    int am[3] = {0, 0, 0};
    if (host_sync_) upload(uh);
    check(mts_error_norms(inst_, eq.t(), &lmax, &l2, am));
    if (!host_sync_) download(uh);
    out << std::setprecision(3) << std::scientific;
    out << "[Xmax,Ymax,Zmax] = [" << am[0] << "," << am[1] << "," << am[2] << "]" << std::endl;
    out << " Lmax = " << lmax << ";  L2 = " << l2 << std::endl;
    out << std::fixed << std::endl;
  }

 protected:
  virtual int method_or_default() const = 0;
  void ensure(int method) {
    if (inst_ && method == method_) return;
    if (inst_) {
      // the LOD variant changed between calls: carry the field over
      std::vector<double> u((size_t)nx_ * ny_ * nz_);
      check(mts_field_download(inst_, u.data(), MTS_LAYOUT_XFAST));
      mts_instance_destroy(inst_);
      inst_ = nullptr;
      check(mts_instance_create(context(), inter_->handle(), &inter_->physics(), method, dt_, &inst_));
      check(mts_field_upload(inst_, u.data(), MTS_LAYOUT_XFAST));
    } else {
      check(mts_instance_create(context(), inter_->handle(), &inter_->physics(), method, dt_, &inst_));
    }
    method_ = method;
  }
  void step(int method, CubicDoub &uh, double t_next) {
    ensure(method);
    if (host_sync_) upload(uh);
    check(mts_step(inst_, t_next));
    if (host_sync_) download(uh);
  }
  void upload(const CubicDoub &uh) {
    if ((int)uh.size() != nx_) fail(MTS_E_ARG, "uh is not sized [nx][ny][nz]");
    flat_.resize((size_t)nx_ * ny_ * nz_);
    for (int i = 0; i < nx_; i++)
      for (int j = 0; j < ny_; j++) {
        if ((int)uh[i].size() != ny_ || (int)uh[i][j].size() != nz_)
          fail(MTS_E_ARG, "uh is not sized [nx][ny][nz]");
        std::memcpy(&flat_[((size_t)i * ny_ + j) * nz_], uh[i][j].data(), nz_ * sizeof(double));
      }
    check(mts_field_upload(inst_, flat_.data(), MTS_LAYOUT_REF));
  }
  void download(CubicDoub &uh) {
    flat_.resize((size_t)nx_ * ny_ * nz_);
    check(mts_field_download(inst_, flat_.data(), MTS_LAYOUT_REF));
    uh.resize(nx_);
    for (int i = 0; i < nx_; i++) {
      uh[i].resize(ny_);
      for (int j = 0; j < ny_; j++) {
        uh[i][j].resize(nz_);
        std::memcpy(uh[i][j].data(), &flat_[((size_t)i * ny_ + j) * nz_], nz_ * sizeof(double));
      }
    }
  }

  Intersections *inter_;
  int nx_, ny_, nz_;
  double dt_ = 0.0;
  mts_instance *inst_ = nullptr;
  int method_ = -1;
  bool host_sync_ = true;
  std::vector<double> flat_;
};
}  // namespace detail

// Douglas-ADI (temporal/douglas_adi.h:47-52); accuracy 2 only.
class Douglas_ADI : public detail::Solver {
 public:
  Douglas_ADI(Intersections &inter, Mesh &mesh, Beta &, VecDoub_I time, Int_I accuracy)
      : Solver(inter, mesh, time) {
    if (accuracy != 2) fail(MTS_E_ARG, "only 2nd-order spatial accuracy is supported");
    ensure(MTS_ADI);
  }
  void Solve_2nd(Equation &eq, Intersections &, CubicDoub_I &uh, Beta &) {
    step(MTS_ADI, uh, eq.t());
  }

 protected:
  int method_or_default() const override { return MTS_ADI; }
};

// LOD, implicit Euler ('I') or Crank-Nicolson ('C') chosen per call (temporal/lod.h:51-56).
class LOD : public detail::Solver {
 public:
  LOD(Intersections &inter, Mesh &mesh, Beta &, VecDoub_I time) : Solver(inter, mesh, time) {}
  void Solve_2nd(Equation &eq, Intersections &, CubicDoub_I &uh, Beta &, char &method) {
    if (method != 'I' && method != 'C') fail(MTS_E_ARG, "LOD method must be 'I' or 'C'");
    step(method == 'I' ? MTS_LOD_IE : MTS_LOD_CN, uh, eq.t());
  }

 protected:
  int method_or_default() const override { return method_ >= 0 ? method_ : MTS_LOD_CN; }
};

// Trapezoidal splitting (temporal/trapezoidal.h:47-53): the three equations
// are at t^n, t^{n+1/2} and t^{n+1}; the library takes them from t^{n+1} and dt.
class TS : public detail::Solver {
 public:
  TS(Intersections &inter, Mesh &mesh, Beta &, VecDoub_I time) : Solver(inter, mesh, time) {
    ensure(MTS_TS);
  }
  void Solve_2nd(Equation &, Equation &, Equation &eq_np1, Intersections &, CubicDoub_I &uh,
                 Beta &) {
    step(MTS_TS, uh, eq_np1.t());
  }

 protected:
  int method_or_default() const override { return MTS_TS; }
};

// --------------------------------------------------------------- LU
// decomposition/lu.h:7-21 over mts_lu_*.
class LU {
 public:
  explicit LU(MatrixDoub_I &a) : n_((int)a.size()) {
    std::vector<double> m((size_t)n_ * n_);
    for (int i = 0; i < n_; i++) {
      if ((int)a[i].size() != n_) fail(MTS_E_ARG, "LU: matrix is not square");
      std::memcpy(&m[(size_t)i * n_], a[i].data(), n_ * sizeof(double));
    }
    mts_lu *l = nullptr;
    check(mts_lu_create(n_, m.data(), &l));
    lu_.reset(l);
  }
  void solve(VecDoub_I &b, VecDoub_O &x) {
    if ((int)b.size() != n_ || (int)x.size() != n_) fail(MTS_E_ARG, "LUdcmp::solve bad sizes");
    check(mts_lu_solve(lu_.get(), b.data(), x.data()));
  }
  void solve(MatrixDoub_I &b, MatrixDoub_O &x) {
    if ((int)b.size() != n_ || (int)x.size() != n_ || b[0].size() != x[0].size())
      fail(MTS_E_ARG, "LUdcmp::solve bad size");
    const int m = (int)b[0].size();
    std::vector<double> bb((size_t)n_ * m), xx((size_t)n_ * m);
    for (int i = 0; i < n_; i++) std::memcpy(&bb[(size_t)i * m], b[i].data(), m * sizeof(double));
    check(mts_lu_solve_matrix(lu_.get(), m, bb.data(), xx.data()));
    for (int i = 0; i < n_; i++) std::memcpy(x[i].data(), &xx[(size_t)i * m], m * sizeof(double));
  }
  void inverse(MatrixDoub_IO &ainv) {
    std::vector<double> m((size_t)n_ * n_);
    check(mts_lu_inverse(lu_.get(), m.data()));
    ainv.assign(n_, VecDoub(n_));
    for (int i = 0; i < n_; i++) std::memcpy(ainv[i].data(), &m[(size_t)i * n_], n_ * sizeof(double));
  }
  double det() {
    double d = 0.0;
    check(mts_lu_det(lu_.get(), &d));
    return d;
  }
  void mprove(VecDoub_I &b, VecDoub_IO &x) {
    if ((int)b.size() != n_ || (int)x.size() != n_) fail(MTS_E_ARG, "LU::mprove bad sizes");
    check(mts_lu_mprove(lu_.get(), b.data(), x.data()));
  }

 private:
  struct Del {
    void operator()(mts_lu *l) const { mts_lu_destroy(l); }
  };
  int n_;
  std::unique_ptr<mts_lu, Del> lu_;
};

}  // namespace mts

#endif  // MTS_HPP
