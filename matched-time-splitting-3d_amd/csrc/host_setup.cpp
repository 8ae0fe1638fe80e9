// Host FP64 setup: grid, classification, crossings, MIB weights.
// See host_setup.hpp for the mapping to the reference.
#include "host_setup.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "host_parallel.hpp"

namespace mts {

static const double kTolSetup = 1.0e-12;   // constant.h:10
static const double kTolRoot = 1.0e-14;    // constant.h:11

Grid make_grid(const double lo[3], const double hi[3], const int n[3]) {
  Grid g;
  for (int a = 0; a < 3; a++) {
    g.n[a] = n[a];
    g.lo[a] = lo[a];
    g.hi[a] = hi[a];
    g.h[a] = (hi[a] - lo[a]) / (n[a] - 1.0);
    g.c[a].resize(n[a]);
    for (int i = 0; i < n[a]; i++) g.c[a][i] = lo[a] + i * g.h[a];
  }
  return g;
}

// ------------------------------------------------------------- surfaces
double Surface::set_node(double x, double y, double z) const {
  return func(x, y, z) + kTolSetup;
}

void Surface::normal(double x, double y, double z, double nrm[3]) const {
  double gx = dfdx(x, y, z), gy = dfdy(x, y, z), gz = dfdz(x, y, z);
  double tot = std::sqrt(gx * gx + gy * gy + gz * gz);
  nrm[0] = gx / tot;
  nrm[1] = gy / tot;
  nrm[2] = gz / tot;
}

double SurfaceEllipsoid::func(double x, double y, double z) const {
  return x * x / (a_ * a_) + y * y / (b_ * b_) + z * z / (c_ * c_) - r_ * r_;
}
double SurfaceEllipsoid::set_node(double x, double y, double z) const {
  return x * x / (a_ * a_) + y * y / (b_ * b_) + z * z / (c_ * c_) - r_ * r_ + kTolSetup;
}
double SurfaceEllipsoid::dfdx(double x, double, double) const { return 2 * x / (a_ * a_); }
double SurfaceEllipsoid::dfdy(double, double y, double) const { return 2 * y / (b_ * b_); }
double SurfaceEllipsoid::dfdz(double, double, double z) const { return 2 * z / (c_ * c_); }
void SurfaceEllipsoid::normal(double x, double y, double z, double nrm[3]) const {
  double tot = std::sqrt(std::pow(2 * x / (a_ * a_), 2) + std::pow(2 * y / (b_ * b_), 2) +
                         std::pow(2 * z / (c_ * c_), 2));
  nrm[0] = 2 * x / (a_ * a_) / tot;
  nrm[1] = 2 * y / (b_ * b_) / tot;
  nrm[2] = 2 * z / (c_ * c_) / tot;
}

double SurfaceTorus::func(double x, double y, double z) const {
  return std::pow((R_ - std::sqrt(x * x + y * y)), 2) + z * z - r_ * r_;
}
double SurfaceTorus::set_node(double x, double y, double z) const {
  return std::pow((R_ - std::sqrt(x * x + y * y)), 2) + z * z - r_ * r_ + kTolSetup;
}
double SurfaceTorus::dfdx(double x, double y, double) const {
  return (-2.0 * R_ * x) / (std::sqrt(x * x + y * y)) + 2.0 * x;
}
double SurfaceTorus::dfdy(double x, double y, double) const {
  return (-2.0 * R_ * y) / (std::sqrt(x * x + y * y)) + 2.0 * y;
}
double SurfaceTorus::dfdz(double, double, double z) const { return 2.0 * z; }
void SurfaceTorus::normal(double x, double y, double z, double nrm[3]) const {
  double tot = std::sqrt(std::pow((-2.0 * R_ * x) / (std::sqrt(x * x + y * y)) + 2.0 * x, 2) +
                         std::pow((-2.0 * R_ * y) / (std::sqrt(x * x + y * y)) + 2.0 * y, 2) +
                         std::pow(2.0 * z, 2));
  nrm[0] = ((-2.0 * R_ * x) / (std::sqrt(x * x + y * y)) + 2.0 * x) / tot;
  nrm[1] = ((-2.0 * R_ * y) / (std::sqrt(x * x + y * y)) + 2.0 * y) / tot;
  nrm[2] = (2.0 * z) / tot;
}

void SurfaceBumpy::grad(double x, double y, double z, double g[3]) const {
  double rho2 = x * x + y * y;
  double r2 = rho2 + z * z;
  double r = std::sqrt(r2);
  double rho = std::sqrt(rho2);
  if (rho == 0.0) {
    g[0] = 0.0;
    g[1] = 0.0;
    g[2] = z / r;
    return;
  }
  double th = std::atan2(rho, z);
  double ps = std::atan2(y, x);
  double dRdth = A_ * m_ * std::cos(m_ * th) * std::cos(k_ * ps);
  double dRdps = -A_ * k_ * std::sin(m_ * th) * std::sin(k_ * ps);
  double dthdx = x * z / (r2 * rho);
  double dthdy = y * z / (r2 * rho);
  double dthdz = -rho / r2;
  double dpsdx = -y / rho2;
  double dpsdy = x / rho2;
  g[0] = x / r - (dRdth * dthdx + dRdps * dpsdx);
  g[1] = y / r - (dRdth * dthdy + dRdps * dpsdy);
  g[2] = z / r - dRdth * dthdz;
}
double SurfaceBumpy::func(double x, double y, double z) const {
  double rho = std::sqrt(x * x + y * y);
  double r = std::sqrt(x * x + y * y + z * z);
  double th = std::atan2(rho, z);
  double ps = std::atan2(y, x);
  return r - (R0_ + A_ * std::sin(m_ * th) * std::cos(k_ * ps));
}
double SurfaceBumpy::dfdx(double x, double y, double z) const {
  double g[3];
  grad(x, y, z, g);
  return g[0];
}
double SurfaceBumpy::dfdy(double x, double y, double z) const {
  double g[3];
  grad(x, y, z, g);
  return g[1];
}
double SurfaceBumpy::dfdz(double x, double y, double z) const {
  double g[3];
  grad(x, y, z, g);
  return g[2];
}
void SurfaceBumpy::normal(double x, double y, double z, double nrm[3]) const {
  double g[3];
  grad(x, y, z, g);
  double tot = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  nrm[0] = g[0] / tot;
  nrm[1] = g[1] / tot;
  nrm[2] = g[2] / tot;
}

double SurfaceCallback::dfdx(double x, double y, double z) const {
  double g[3];
  grad_(user_, x, y, z, g);
  return g[0];
}
double SurfaceCallback::dfdy(double x, double y, double z) const {
  double g[3];
  grad_(user_, x, y, z, g);
  return g[1];
}
double SurfaceCallback::dfdz(double x, double y, double z) const {
  double g[3];
  grad_(user_, x, y, z, g);
  return g[2];
}

// ---------------------------------------------------------------- beta
double BetaModel::inside(double x, double y, double) const {
  if (kind == 1) return p[0] * (1.0 + 0.5 * std::sin(x) * std::cos(y));
  return p[0];
}
double BetaModel::outside(double, double, double) const { return p[1]; }

// -------------------------------------------------------------- dense LU
DenseLU::DenseLU(int n, const double *a) : n_(n), lu_(a, a + (size_t)n * n), perm_(n) {
  const double tiny = 1.0e-40;
  std::vector<double> scale(n);
  for (int i = 0; i < n; i++) {
    double big = 0.0;
    for (int j = 0; j < n; j++) big = std::fmax(big, std::fabs(lu_[(size_t)i * n + j]));
    if (big == 0.0) {
      singular_row_ = true;
      big = 1.0;
    }
    scale[i] = 1.0 / big;
  }
  // The pivot row index persists across columns, as in the textbook routine:
  // an all-zero column keeps the previous pivot choice.
  int piv = 0;
  for (int k = 0; k < n; k++) {
    double big = 0.0;
    for (int i = k; i < n; i++) {
      double t = scale[i] * std::fabs(lu_[(size_t)i * n + k]);
      if (t > big) {
        big = t;
        piv = i;
      }
    }
    if (piv != k) {
      for (int j = 0; j < n; j++) std::swap(lu_[(size_t)piv * n + j], lu_[(size_t)k * n + j]);
      sign_ = -sign_;
      scale[piv] = scale[k];
    }
    perm_[k] = piv;
    double &pkk = lu_[(size_t)k * n + k];
    if (pkk == 0.0) pkk = tiny;
    for (int i = k + 1; i < n; i++) {
      double m = lu_[(size_t)i * n + k] /= lu_[(size_t)k * n + k];
      for (int j = k + 1; j < n; j++) lu_[(size_t)i * n + j] -= m * lu_[(size_t)k * n + j];
    }
  }
}

void DenseLU::solve(const double *b, double *x) const {
  const int n = n_;
  if (x != b)
    for (int i = 0; i < n; i++) x[i] = b[i];
  int first = 0;  // 1 + index of the first nonzero of the permuted b
  for (int i = 0; i < n; i++) {
    int ip = perm_[i];
    double s = x[ip];
    x[ip] = x[i];
    if (first != 0) {
      for (int j = first - 1; j < i; j++) s -= lu_[(size_t)i * n + j] * x[j];
    } else if (s != 0.0) {
      first = i + 1;
    }
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    for (int j = i + 1; j < n; j++) s -= lu_[(size_t)i * n + j] * x[j];
    x[i] = s / lu_[(size_t)i * n + i];
  }
}

void DenseLU::inverse(double *ainv) const {
  const int n = n_;
  std::vector<double> col(n);
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < n; i++) col[i] = (i == j) ? 1.0 : 0.0;
    solve(col.data(), col.data());
    for (int i = 0; i < n; i++) ainv[(size_t)i * n + j] = col[i];
  }
}

double DenseLU::det() const {
  double d = sign_;
  for (int i = 0; i < n_; i++) d *= lu_[(size_t)i * n_ + i];
  return d;
}

// --------------------------------------------------------------- Fornberg
void fornberg_weights(double z, const double *x, int n, int m, double *c) {
  const int w = m + 1;
  for (int i = 0; i < n * w; i++) c[i] = 0.0;
  double c1 = 1.0;
  double c4 = x[0] - z;
  c[0] = 1.0;
  for (int i = 1; i < n; i++) {
    int mn = i < m ? i : m;
    double c2 = 1.0;
    double c5 = c4;
    c4 = x[i] - z;
    for (int j = 0; j < i; j++) {
      double c3 = x[i] - x[j];
      c2 = c2 * c3;
      if (j == i - 1) {
        for (int k = mn; k > 0; k--)
          c[i * w + k] = c1 * (k * c[(i - 1) * w + k - 1] - c5 * c[(i - 1) * w + k]) / c2;
        c[i * w + 0] = -c1 * c5 * c[(i - 1) * w + 0] / c2;
      }
      for (int k = mn; k > 0; k--) c[j * w + k] = (c4 * c[j * w + k] - k * c[j * w + k - 1]) / c3;
      c[j * w + 0] = c4 * c[j * w + 0] / c3;
    }
    c1 = c2;
  }
}

// ------------------------------------------------------------- crossings
int SetupResult::line_nj(int d) const { return d == 0 ? grid.n[1] : grid.n[0]; }
int SetupResult::line_id(int d, int j, int k) const { return k * line_nj(d) + j; }

namespace {

// axis layout of a line in direction d: the line coordinate is axis d, the
// fixed indices (j, k) are axes (ja, ka).
inline void line_axes(int d, int &ja, int &ka) {
  if (d == 0) { ja = 1; ka = 2; }
  else if (d == 1) { ja = 0; ka = 2; }
  else { ja = 0; ka = 1; }
}

// Safeguarded Newton-bisection for a root of phi along axis d between p1 and
// p2 (Press et al., Numerical Recipes "rtsafe"), stopping when the last step
// is below 1e-14: the iteration the reference uses (surface_cartesian.cc:134-373).
bool root_on_axis(const Surface &s, int d, const double base[3], double p1, double p2,
                  double &root, std::string &err) {
  auto f_at = [&](double v) {
    double q[3] = {base[0], base[1], base[2]};
    q[d] = v;
    return s.func(q[0], q[1], q[2]);
  };
  auto df_at = [&](double v) {
    double q[3] = {base[0], base[1], base[2]};
    q[d] = v;
    return s.dfd(d, q[0], q[1], q[2]);
  };
  if (p1 == p2) {
    err = "two bounds are too close";
    return false;
  }
  if (std::fabs(p1 - p2) < kTolRoot) {
    root = (p1 + p2) / 2;
    return true;
  }
  double fl = f_at(p1), fh = f_at(p2);
  if ((fl > 0.0 && fh > 0.0) || (fl < 0.0 && fh < 0.0)) {
    err = "root must be bracketed";
    return false;
  }
  if (fl == 0.0) { root = p1; return true; }
  if (fh == 0.0) { root = p2; return true; }
  double lo_pt, hi_pt;  // phi(lo_pt) < 0 < phi(hi_pt)
  if (fl < 0.0) { lo_pt = p1; hi_pt = p2; }
  else { hi_pt = p1; lo_pt = p2; }
  double r = 0.5 * (p1 + p2);
  double step_prev = std::fabs(p2 - p1);
  double step = step_prev;
  double f = f_at(r), df = df_at(r);
  for (int it = 0; it < 100; it++) {
    bool newton_out = (((r - hi_pt) * df - f) * ((r - lo_pt) * df - f) > 0.0);
    bool slow = (std::fabs(2.0 * f) > std::fabs(step_prev * df));
    if (newton_out || slow) {
      step_prev = step;
      step = 0.5 * (hi_pt - lo_pt);
      r = lo_pt + step;
      if (lo_pt == r) { root = r; return true; }
    } else {
      step_prev = step;
      step = f / df;
      double prev = r;
      r -= step;
      if (prev == r) { root = r; return true; }
    }
    if (std::fabs(step) < kTolRoot) { root = r; return true; }
    f = f_at(r);
    df = df_at(r);
    if (f < 0.0) lo_pt = r;
    else hi_pt = r;
  }
  err = "maximum number of root iterations exceeded";
  return false;
}

// Local frame rows (n, eta, tau) for a crossing on a d-line
// (Pmatrix_Setup_x/y/z, intersection.cc:2503-2598).
void local_frame(int d, const double a[3], double p[3][3]) {
  for (int i = 0; i < 3; i++) p[0][i] = a[i];
  if (d == 0) {
    p[1][0] = a[2] / std::sqrt(a[0] * a[0] + a[2] * a[2]);
    p[1][1] = 0;
    p[1][2] = -a[0] / std::sqrt(a[0] * a[0] + a[2] * a[2]);
    p[2][0] = a[1] / std::sqrt(a[0] * a[0] + a[1] * a[1]);
    p[2][1] = -a[0] / std::sqrt(a[0] * a[0] + a[1] * a[1]);
    p[2][2] = 0;
  } else if (d == 1) {
    p[1][0] = 0;
    p[1][1] = a[2] / std::sqrt(a[1] * a[1] + a[2] * a[2]);
    p[1][2] = -a[1] / std::sqrt(a[1] * a[1] + a[2] * a[2]);
    p[2][0] = a[1] / std::sqrt(a[0] * a[0] + a[1] * a[1]);
    p[2][1] = -a[0] / std::sqrt(a[0] * a[0] + a[1] * a[1]);
    p[2][2] = 0;
  } else {
    p[1][0] = 0;
    p[1][1] = a[2] / std::sqrt(a[1] * a[1] + a[2] * a[2]);
    p[1][2] = -a[1] / std::sqrt(a[1] * a[1] + a[2] * a[2]);
    p[2][0] = a[2] / std::sqrt(a[0] * a[0] + a[2] * a[2]);
    p[2][1] = 0;
    p[2][2] = -a[0] / std::sqrt(a[0] * a[0] + a[2] * a[2]);
  }
}

// Crossings of a range of lines: reads the classification, appends to its
// own crossing list (one Builder per chunk of lines).
struct Builder {
  const Grid &g;
  const Surface &s;
  const BetaModel &beta;
  const SetupResult &out;          // mib and ind
  std::vector<Crossing> &cross;    // this chunk's crossings, in line order
  int singular_warnings = 0;
  int corner_pairs = 0;
  std::string err;

  // Geometry of one crossing in the cell (i, i+1) of line (d, j, k).
  bool fill(int d, int j, int k, int i, int id, Crossing &c) {
    int ja, ka;
    line_axes(d, ja, ka);
    std::memset(&c, 0, sizeof(c));
    c.dir = d;
    c.j = j;
    c.k = k;
    c.id = id;
    c.left_loc = i;
    double base[3];
    base[ja] = g.c[ja][j];
    base[ka] = g.c[ka][k];
    base[d] = 0.0;
    double root;
    if (!root_on_axis(s, d, base, g.c[d][i], g.c[d][i + 1], root, err)) return false;
    c.coord[ja] = base[ja];
    c.coord[ka] = base[ka];
    c.coord[d] = root;
    c.gamma = root - g.c[d][i];
    c.beta_in = beta.inside(c.coord[0], c.coord[1], c.coord[2]);
    c.beta_out = beta.outside(c.coord[0], c.coord[1], c.coord[2]);
    s.normal(c.coord[0], c.coord[1], c.coord[2], c.normal);
    local_frame(d, c.normal, c.p);
    double pm[9], pinv[9];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) pm[a * 3 + b] = c.p[a][b];
    DenseLU lu(3, pm);
    lu.inverse(pinv);
    c.pinv_d0 = pinv[d * 3 + 0];
    return true;
  }

  void warn_if_singular(const DenseLU &lu) {
    if (std::fabs(lu.det()) < kTolSetup) singular_warnings++;
  }

  // MIB-L1 pair of fictitious points from 2-point interpolations
  // (intersection.cc:930-1066): x[0..3] weights of the FP at left_loc,
  // x[4..7] weights of the FP at left_loc+1, over (u_i, u_{i+1}, [u], [beta u_d]).
  void mib_l1(const Crossing &c, double dv, double x[8]) {
    double v[2] = {0.0, dv};
    double t[4];
    fornberg_weights(c.gamma, v, 2, 1, t);
    double w0[2] = {t[0 * 2 + 0], t[1 * 2 + 0]};
    double w1[2] = {t[0 * 2 + 1], t[1 * 2 + 1]};
    double A[64] = {0}, B[8] = {0};
    bool even = (std::abs(c.id) % 2 == 0);
    B[0] = -w0[0];
    B[1] = w0[1];
    for (int i = 0; i < 4; i++) {
      A[i * 8 + i] = -w0[0];
      A[i * 8 + 4 + i] = w0[1];
    }
    B[2] += even ? -1 : 1;
    double coefr = even ? c.beta_in : c.beta_out;
    double coefl = even ? c.beta_out : c.beta_in;
    B[4] = -w1[0] * coefr;
    B[5] = w1[1] * coefl;
    for (int i = 0; i < 4; i++) {
      A[(4 + i) * 8 + i] = -w1[0] * coefl;
      A[(4 + i) * 8 + 4 + i] = w1[1] * coefr;
    }
    B[7] += even ? -1 : 1;
    DenseLU lu(8, A);
    lu.solve(B, x);
    warn_if_singular(lu);
  }

  // MIB-L2 irregular crossing, 2nd order (intersection.cc:1209-1395,
  // Get_irr_weights :1424-1468): 12 unknowns, weights over
  // (u_{i-1}, u_i, u_{i+1}, u_{i+2}, [u], [beta u_d]).
  void mib_l2_irregular(Crossing &c, double dv) {
    const int nu = 6, ns = 3, side = 2;
    double vr[3], vl[3];
    for (int i = 0; i < ns; i++) {
      vr[i] = -(side - 1) * dv + i * dv;
      vl[i] = (side - ns + 1) * dv + i * dv;
    }
    double tr[6], tl[6];
    fornberg_weights(c.gamma, vr, ns, 1, tr);
    fornberg_weights(c.gamma, vl, ns, 1, tl);
    auto WR = [&](int o, int i) { return tr[i * 2 + o]; };
    auto WL = [&](int o, int i) { return tl[i * 2 + o]; };
    double A[144] = {0}, B[12] = {0};
    bool even = (std::abs(c.id) % 2 == 0);
    for (int i = 0; i < side; i++) {
      B[i] += -WR(0, i);
      B[side + i] += WL(0, ns - side + i);
    }
    for (int i = 0; i < nu; i++) {
      A[i * 12 + i] = -WL(0, 0);
      A[i * 12 + nu + i] = WR(0, ns - 1);
    }
    B[nu - 2] += even ? -1 : 1;
    double coefr = even ? c.beta_in : c.beta_out;
    double coefl = even ? c.beta_out : c.beta_in;
    for (int i = 0; i < side; i++) {
      B[nu + i] += -WR(1, i) * coefr;
      B[nu + side + i] += WL(1, ns - side + i) * coefl;
    }
    for (int i = 0; i < nu; i++) {
      A[(nu + i) * 12 + nu + i] = WR(1, ns - 1) * coefr;
      A[(nu + i) * 12 + i] = -WL(1, 0) * coefl;
    }
    B[2 * nu - 1] += even ? -1 : 1;
    DenseLU lu(12, A);
    double x[12];
    lu.solve(B, x);
    warn_if_singular(lu);
    c.nw = 6;
    for (int j = 0; j < nu; j++) {
      c.wl[j] = x[j];
      c.wr[j] = x[nu + j];
    }
  }

  // MIB-L2 corner pair, 2nd order (Corner_MIB_L2_2nd intersection.cc:1479-1704,
  // Get_cor_weights :1752-1830): four FPs x 9 components solved together.
  void mib_l2_corner(Crossing &L, Crossing &R, double dv) {
    const int nu = 9, ns = 4;
    double lvo[4], lvi[4], rvo[4], rvi[4];
    for (int i = 0; i < ns; i++) {
      lvo[i] = -dv + i * dv;
      rvo[i] = i * dv;
    }
    for (int i = 0; i < ns; i++) lvi[i] = i * dv;
    if ((dv - L.gamma) > R.gamma) lvi[ns - 1] = -dv;
    for (int i = 0; i < ns; i++) rvi[i] = lvi[i];
    double to_l[8], ti_l[8], to_r[8], ti_r[8];
    fornberg_weights(L.gamma, lvo, ns, 1, to_l);
    fornberg_weights(L.gamma, lvi, ns, 1, ti_l);
    fornberg_weights(dv + R.gamma, rvo, ns, 1, to_r);
    fornberg_weights(dv + R.gamma, rvi, ns, 1, ti_r);
    auto LO = [&](int o, int i) { return to_l[i * 2 + o]; };
    auto LI = [&](int o, int i) { return ti_l[i * 2 + o]; };
    auto RO = [&](int o, int i) { return to_r[i * 2 + o]; };
    auto RI = [&](int o, int i) { return ti_r[i * 2 + o]; };
    const int N = 4 * nu;
    std::vector<double> A((size_t)N * N, 0.0);
    double B[36] = {0};
    auto a = [&](int r, int col) -> double & { return A[(size_t)r * N + col]; };
    bool even = (std::abs(L.id) % 2 == 0);
    // left interface, value
    B[0] = -LO(0, 0);
    B[1] = -LO(0, 1);
    B[2] = LI(0, 1);
    B[3] = -LO(0, 3);
    for (int i = 0; i < nu; i++) {
      a(i, i) = -LI(0, 0);
      a(i, nu + i) = LO(0, 2);
      a(i, 2 * nu + i) = -LI(0, 2);
      a(i, 3 * nu + i) = -LI(0, 3);
    }
    B[5] = even ? -1 : 1;
    // left interface, flux
    double cin = even ? L.beta_out : L.beta_in;
    double cout_ = even ? L.beta_in : L.beta_out;
    B[nu + 0] = -LO(1, 0) * cout_;
    B[nu + 1] = -LO(1, 1) * cout_;
    B[nu + 2] = LI(1, 1) * cin;
    B[nu + 3] = -LO(1, 3) * cout_;
    for (int i = 0; i < nu; i++) {
      a(nu + i, i) = -LI(1, 0) * cin;
      a(nu + i, nu + i) = LO(1, 2) * cout_;
      a(nu + i, 2 * nu + i) = -LI(1, 2) * cin;
      a(nu + i, 3 * nu + i) = -LI(1, 3) * cin;
    }
    B[nu + 6] = even ? -1 : 1;
    // right interface, value
    B[2 * nu + 1] = -RO(0, 0);
    B[2 * nu + 2] = RI(0, 1);
    B[2 * nu + 3] = -RO(0, 2);
    B[2 * nu + 4] = -RO(0, 3);
    for (int i = 0; i < nu; i++) {
      a(2 * nu + i, i) = -RI(0, 0);
      a(2 * nu + i, nu + i) = RO(0, 1);
      a(2 * nu + i, 2 * nu + i) = -RI(0, 2);
      a(2 * nu + i, 3 * nu + i) = -RI(0, 3);
    }
    B[2 * nu + 7] = even ? -1 : 1;
    // right interface, flux (sides fixed by the left crossing's parity)
    cin = even ? R.beta_out : R.beta_in;
    cout_ = even ? R.beta_in : R.beta_out;
    B[3 * nu + 1] = -RO(1, 0) * cout_;
    B[3 * nu + 2] = RI(1, 1) * cin;
    B[3 * nu + 3] = -RO(1, 2) * cout_;
    B[3 * nu + 4] = -RO(1, 3) * cout_;
    for (int i = 0; i < nu; i++) {
      a(3 * nu + i, i) = -RI(1, 0) * cin;
      a(3 * nu + i, nu + i) = RO(1, 1) * cout_;
      a(3 * nu + i, 2 * nu + i) = -RI(1, 2) * cin;
      a(3 * nu + i, 3 * nu + i) = -RI(1, 3) * cin;
    }
    B[3 * nu + 8] = even ? -1 : 1;
    DenseLU lu(N, A.data());
    double x[36];
    lu.solve(B, x);
    warn_if_singular(lu);
    L.nw = R.nw = 9;
    for (int j = 0; j < nu; j++) {
      L.wl[j] = x[j];
      L.wr[j] = x[nu + j];
      R.wl[j] = x[nu + j];
      R.wr[j] = x[2 * nu + j];
    }
  }

  void mib_l1_irregular(Crossing &c, double dv) {
    double x[8];
    mib_l1(c, dv, x);
    c.nw = 6;
    std::memset(c.wl, 0, sizeof(c.wl));
    std::memset(c.wr, 0, sizeof(c.wr));
    c.wl[1] = x[0]; c.wl[2] = x[1]; c.wl[4] = x[2]; c.wl[5] = x[3];
    c.wr[1] = x[4]; c.wr[2] = x[5]; c.wr[4] = x[6]; c.wr[5] = x[7];
  }

  void mib_l1_corner(Crossing &L, Crossing &R, double dv) {
    double x[8];
    L.nw = R.nw = 9;
    std::memset(L.wl, 0, sizeof(L.wl));
    std::memset(L.wr, 0, sizeof(L.wr));
    std::memset(R.wl, 0, sizeof(R.wl));
    std::memset(R.wr, 0, sizeof(R.wr));
    mib_l1(L, dv, x);
    L.wl[1] = x[0]; L.wl[2] = x[1]; L.wl[5] = x[2]; L.wl[6] = x[3];
    L.wr[1] = x[4]; L.wr[2] = x[5]; L.wr[5] = x[6]; L.wr[6] = x[7];
    mib_l1(R, dv, x);
    R.wl[2] = x[0]; R.wl[3] = x[1]; R.wl[7] = x[2]; R.wl[8] = x[3];
    R.wr[2] = x[4]; R.wr[3] = x[5]; R.wr[7] = x[6]; R.wr[8] = x[7];
  }

  // Scan one line for sign changes (Setup_Intersections intersection.cc:79-291).
  bool scan_line(int d, int j, int k) {
    int ja, ka;
    line_axes(d, ja, ka);
    const int n = g.n[d];
    auto sgn = [&](int i) -> int {
      int q[3];
      q[ja] = j;
      q[ka] = k;
      q[d] = i;
      return out.ind[g.idx(q[0], q[1], q[2])];
    };
    static const char *names = "xyz";
    char where[96];
    std::snprintf(where, sizeof(where), " on the %c-line (%d, %d)", names[d], j, k);
    if (sgn(0) * sgn(1) < 0) {
      err = std::string("interface crosses the first cell of a line") + where;
      return false;
    }
    int count = 0;
    int i = 1;
    while (i < n - 1) {
      if (sgn(i) * sgn(i + 1) < 0) {
        if (i + 2 > n - 1) {
          err = std::string("interface crosses the last cell of a line") + where;
          return false;
        }
        if (sgn(i + 1) * sgn(i + 2) > 0) {
          Crossing c;
          count += 1;
          if (!fill(d, j, k, i, count, c)) return false;
          if (out.mib == 1) mib_l1_irregular(c, g.h[d]);
          else mib_l2_irregular(c, g.h[d]);
          cross.push_back(c);
          i += 2;
        } else {
          if (i + 3 > n - 1) {
            err = std::string("corner interface touches the boundary") + where;
            return false;
          }
          if (sgn(i + 2) * sgn(i + 3) < 0) {
            err = std::string("two corner interfaces coherent") + where;
            return false;
          }
          Crossing L, R;
          count += 1;
          if (!fill(d, j, k, i, -count, L)) return false;
          count += 1;
          if (!fill(d, j, k, i + 1, -count, R)) return false;
          if (out.mib == 1) mib_l1_corner(L, R, g.h[d]);
          else mib_l2_corner(L, R, g.h[d]);
          cross.push_back(L);
          cross.push_back(R);
          corner_pairs++;
          i += 3;
        }
      } else {
        i += 1;
      }
    }
    if (count % 2 == 1) {
      err = std::string("open surface") + where;
      return false;
    }
    return true;
  }
};

}  // namespace

std::string classify_nodes(const Grid &g, const Surface &s, std::vector<int8_t> &ind) {
  // mesh.cc:51-69: set_node > 1e-12 outside, < 1e-12 inside, equal: exit.
  // Planes of z in parallel; the error named is the first node in the
  // reference's order that lies on the interface.
  ind.resize(g.nodes());
  const int nz = g.n[2];
  const int threads = s.concurrent() ? setup_threads() : 1;
  const int chunks = chunk_count(nz, threads);
  std::vector<std::string> errs((size_t)chunks);
  parallel_chunks(nz, chunks, threads, [&](int c, int z0, int z1) {
    for (int iz = z0; iz < z1; iz++)
      for (int iy = 0; iy < g.n[1]; iy++)
        for (int ix = 0; ix < g.n[0]; ix++) {
          double v = s.set_node(g.c[0][ix], g.c[1][iy], g.c[2][iz]);
          int8_t t;
          if (v > kTolSetup) t = 1;
          else if (v < kTolSetup) t = -1;
          else {
            char buf[128];
            std::snprintf(buf, sizeof(buf), "interface on grid node (%d, %d, %d)", ix, iy, iz);
            errs[(size_t)c] = buf;
            return;
          }
          ind[g.idx(ix, iy, iz)] = t;
        }
  });
  for (auto &e : errs)
    if (!e.empty()) return e;
  return std::string();
}

std::string build_setup(const Grid &g, const Surface &s, const BetaModel &beta,
                        const BuildOptions &opt, SetupResult &out) {
  out = SetupResult();
  out.grid = g;
  out.mib = opt.mib;
  out.accuracy = opt.accuracy;
  if (opt.mib != 1 && opt.mib != 2) return "MIB method must be 1 (L1) or 2 (L2)";
  if (opt.accuracy != 2) return "only 2nd-order spatial accuracy is supported";
  for (int a = 0; a < 3; a++)
    if (g.n[a] < 5) return "each axis needs at least 5 nodes";
  {
    std::string err = classify_nodes(g, s, out.ind);
    if (!err.empty()) return err;
  }
  // Lines of one direction in chunks of rows k, each chunk its own crossing
  // list; stitched in line order, so crossings, line_ptr and the error named
  // (the first line in the reference's scan order) match a one-thread scan.
  const int threads = s.concurrent() ? setup_threads() : 1;
  for (int d = 0; d < 3; d++) {
    int ja, ka;
    line_axes(d, ja, ka);
    int nj = g.n[ja], nk = g.n[ka];
    out.n_lines[d] = nj * nk;
    out.line_ptr[d].assign((size_t)nj * nk + 1, 0);
    const int chunks = chunk_count(nk, threads);
    std::vector<std::vector<Crossing>> part((size_t)chunks);
    std::vector<Builder> bs;
    bs.reserve((size_t)chunks);
    for (int c = 0; c < chunks; c++) bs.push_back(Builder{g, s, beta, out, part[(size_t)c], 0, 0, std::string()});
    parallel_chunks(nk, chunks, threads, [&](int c, int k0, int k1) {
      Builder &b = bs[(size_t)c];
      for (int k = k0; k < k1; k++)
        for (int j = 0; j < nj; j++) {
          int lid = k * nj + j;
          out.line_ptr[d][lid] = (int)b.cross.size();  // chunk-local until stitched
          if (j == 0 || k == 0 || j == nj - 1 || k == nk - 1) continue;
          if (!b.scan_line(d, j, k)) return;
        }
    });
    for (auto &b : bs)
      if (!b.err.empty()) return b.err;
    size_t total = 0;
    for (auto &pc : part) total += pc.size();
    if (total > (size_t)INT32_MAX) return "too many interface crossings";
    out.cross[d].reserve(total);
    for (int c = 0; c < chunks; c++) {
      const int base = (int)out.cross[d].size();
      const int l0 = chunk_begin(nk, chunks, c) * nj, l1 = chunk_begin(nk, chunks, c + 1) * nj;
      for (int lid = l0; lid < l1; lid++) out.line_ptr[d][lid] += base;
      out.cross[d].insert(out.cross[d].end(), part[(size_t)c].begin(), part[(size_t)c].end());
      std::vector<Crossing>().swap(part[(size_t)c]);
      out.singular_warnings += bs[(size_t)c].singular_warnings;
      out.corner_pairs[d] += bs[(size_t)c].corner_pairs;
    }
    out.line_ptr[d][(size_t)nj * nk] = (int)out.cross[d].size();
  }
  return "";
}

}  // namespace mts
