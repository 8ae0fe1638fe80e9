// Host-side FP64 setup of the Matched-TS hot path: grid, node classification,
// interface crossings per grid line and the MIB fictitious-point weights.
//
// This is the one-time part of the path (reference: Mesh::Mesh
// mesh/mesh.cc:20-70, Intersections::Intersections spatial/intersection.cc:22-917,
// LU decomposition/lu.cc, Surface_Cartesian::Rtsafe_* surface/surface_cartesian.cc).
// It stays on the host in FP64 and produces the sparse table the device
// kernels consume every step.  All arithmetic that feeds a device-visible
// value is evaluated in the same operation order as the reference so the
// table matches it to the last bit where the inputs do.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace mts {

// ---------------------------------------------------------------- grid
struct Grid {
  int n[3];                    // node counts per axis
  double lo[3], hi[3];         // domain bounds
  double h[3];                 // spacing, (hi - lo) / (n - 1.0)  (mesh.cc:35-37)
  std::vector<double> c[3];    // node coordinates, lo + i*h       (mesh.cc:39-49)
  size_t nodes() const { return (size_t)n[0] * n[1] * n[2]; }
  // x-fastest linear index, the reference's To1d (mesh.cc:93-99)
  size_t idx(int ix, int iy, int iz) const {
    return ((size_t)iz * n[1] + iy) * n[0] + ix;
  }
};
Grid make_grid(const double lo[3], const double hi[3], const int n[3]);

// ------------------------------------------------------------- surfaces
// A level set phi: phi > 0 outside (Omega+), phi < 0 inside (Omega-).
// Mirrors the reference's Surface_Cartesian (surface_cartesian.h:6-26):
// set_node = phi + 1e-12 (classification), func, its partial derivatives and
// the unit normal.
class Surface {
 public:
  virtual ~Surface() = default;
  virtual double func(double x, double y, double z) const = 0;
  virtual double dfdx(double x, double y, double z) const = 0;
  virtual double dfdy(double x, double y, double z) const = 0;
  virtual double dfdz(double x, double y, double z) const = 0;
  virtual double set_node(double x, double y, double z) const;
  virtual void normal(double x, double y, double z, double nrm[3]) const;
  // Whether several threads may evaluate the surface at once (the built-in
  // shapes; not a caller's callbacks, which run on one thread).
  virtual bool concurrent() const { return true; }
  double dfd(int axis, double x, double y, double z) const {
    return axis == 0 ? dfdx(x, y, z) : axis == 1 ? dfdy(x, y, z) : dfdz(x, y, z);
  }
};

// x^2/a^2 + y^2/b^2 + z^2/c^2 - r^2 (surface/surface_ellipsoid.cc)
class SurfaceEllipsoid : public Surface {
  double a_, b_, c_, r_;

 public:
  SurfaceEllipsoid(double a, double b, double c, double r) : a_(a), b_(b), c_(c), r_(r) {}
  double func(double x, double y, double z) const override;
  double dfdx(double x, double y, double z) const override;
  double dfdy(double x, double y, double z) const override;
  double dfdz(double x, double y, double z) const override;
  double set_node(double x, double y, double z) const override;
  void normal(double x, double y, double z, double nrm[3]) const override;
};

// (R - sqrt(x^2 + y^2))^2 + z^2 - r^2 (surface/surface_torus.cc)
class SurfaceTorus : public Surface {
  double R_, r_;

 public:
  SurfaceTorus(double R, double r) : R_(R), r_(r) {}
  double func(double x, double y, double z) const override;
  double dfdx(double x, double y, double z) const override;
  double dfdy(double x, double y, double z) const override;
  double dfdz(double x, double y, double z) const override;
  double set_node(double x, double y, double z) const override;
  void normal(double x, double y, double z, double nrm[3]) const override;
};

// |x| - (R0 + A sin(m theta) cos(k psi)), theta = atan2(rho, z), psi = atan2(y, x).
// On the z axis the gradient is taken radial (the psi terms have no limit).
class SurfaceBumpy : public Surface {
  double R0_, A_, m_, k_;
  void grad(double x, double y, double z, double g[3]) const;

 public:
  SurfaceBumpy(double R0, double A, double m, double k) : R0_(R0), A_(A), m_(m), k_(k) {}
  double func(double x, double y, double z) const override;
  double dfdx(double x, double y, double z) const override;
  double dfdy(double x, double y, double z) const override;
  double dfdz(double x, double y, double z) const override;
  void normal(double x, double y, double z, double nrm[3]) const override;
};

// User level set through C callbacks (C-ABI surfaces).
class SurfaceCallback : public Surface {
  double (*phi_)(void *, double, double, double);
  void (*grad_)(void *, double, double, double, double *);
  void *user_;

 public:
  SurfaceCallback(double (*phi)(void *, double, double, double),
                  void (*grad)(void *, double, double, double, double *), void *user)
      : phi_(phi), grad_(grad), user_(user) {}
  double func(double x, double y, double z) const override { return phi_(user_, x, y, z); }
  double dfdx(double x, double y, double z) const override;
  double dfdy(double x, double y, double z) const override;
  double dfdz(double x, double y, double z) const override;
  bool concurrent() const override { return false; }
};

// -------------------------------------------------------- diffusion beta
// kind 0: beta- = p0, beta+ = p1
// kind 1: beta- = p0 (1 + 0.5 sin x cos y), beta+ = p1
struct BetaModel {
  int kind = 0;
  double p[4] = {1.0, 1.0, 0.0, 0.0};
  double inside(double x, double y, double z) const;
  double outside(double x, double y, double z) const;
};

// ----------------------------------------------------------- dense LU
// Crout LU with implicit-scaled partial pivoting (Press et al., Numerical
// Recipes, ludcmp/lubksb), the decomposition the reference uses for every
// small dense system (decomposition/lu.cc:15-115).  Row-major storage.
class DenseLU {
 public:
  explicit DenseLU(int n, const double *a);  // a: n*n row-major
  void solve(const double *b, double *x) const;  // b may alias x
  void inverse(double *ainv) const;              // n*n row-major
  double det() const;
  int n() const { return n_; }
  const std::vector<double> &factors() const { return lu_; }
  const std::vector<int> &perm() const { return perm_; }
  bool singular_row() const { return singular_row_; }

 private:
  int n_;
  std::vector<double> lu_;
  std::vector<int> perm_;
  double sign_ = 1.0;
  bool singular_row_ = false;
};

// Fornberg's recursion for finite-difference weights (B. Fornberg, Math.
// Comp. 51 (1988) 699-706): weights c[i*(m+1) + k] of node x[i] for the k-th
// derivative at z, k = 0..m.
void fornberg_weights(double z, const double *x, int n, int m, double *c);

// -------------------------------------------------------------- crossings
// One interface crossing on a grid line (the live part of the reference's
// Intersection_Data, constant.h:141-156).
struct Crossing {
  int dir;          // 0 = x, 1 = y, 2 = z line
  int j, k;         // the line's two fixed indices (x: iy,iz; y: ix,iz; z: ix,iy)
  int id;           // 1-based count along the line; negative for a corner pair
  int left_loc;     // index of the node left of the crossing
  double gamma;     // crossing position - coordinate of left_loc
  double coord[3];  // crossing point
  double normal[3];
  double p[3][3];   // local frame (n, eta, tau), Pmatrix_Setup_* (intersection.cc:2503-2598)
  double pinv_d0;   // P^{-1}[dir][0]: u_dir = pinv_d0 * [beta u_xi]
  double beta_in, beta_out;
  int nw;           // 6 (irregular) or 9 (corner)
  double wl[9];     // weil[0]: FP of the right side at left_loc
  double wr[9];     // weir[0]: FP of the left side at left_loc + 1
};

struct SetupResult {
  Grid grid;
  int mib = 1;
  int accuracy = 2;
  std::vector<int8_t> ind;                  // +1 outside, -1 inside, x-fastest
  std::vector<Crossing> cross[3];           // per direction, in line-id order
  std::vector<int> line_ptr[3];             // CSR: line id -> cross[] range
  int n_lines[3] = {0, 0, 0};
  int singular_warnings = 0;                // MIB systems with |det| < 1e-12
  int corner_pairs[3] = {0, 0, 0};
  // line id for direction d, fixed indices (j, k)
  int line_id(int d, int j, int k) const;
  int line_nj(int d) const;
};

struct BuildOptions {
  int mib = 1;       // 1 = MIB-L1, 2 = MIB-L2
  int accuracy = 2;  // only 2nd order is on the hot path
};

// Node classification of Mesh::Mesh (mesh/mesh.cc:51-69): +1 outside, -1
// inside, x-fastest.  Returns "" or the node the interface passes through.
std::string classify_nodes(const Grid &g, const Surface &s, std::vector<int8_t> &ind);

// Returns "" on success, else the reason (the reference would have exited).
std::string build_setup(const Grid &g, const Surface &s, const BetaModel &beta,
                        const BuildOptions &opt, SetupResult &out);

}  // namespace mts
