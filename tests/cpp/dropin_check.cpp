// Checks of the C++ drop-in classes (include/mts.hpp), printed as one JSON
// line per check for tests/test_dropin.py and tests/test_gpu_dropin.py.
//   dropin_check host   LU, Mesh, Intersections and error modes (no GPU)
//   dropin_check gpu    the solver classes on a small problem
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "mts.hpp"

using namespace mts;

static void emit(const char *name, bool ok, double value) {
  std::printf("{\"check\": \"%s\", \"ok\": %s, \"value\": %.17g}\n", name, ok ? "true" : "false",
              value);
  std::fflush(stdout);
}

static int host_checks() {
  // LU on a 4x4 system with a known solution
  MatrixDoub a = {{4, -2, 1, 0.5}, {-2, 5, -1, 2}, {1, -1, 6, -3}, {0.5, 2, -3, 7}};
  VecDoub x_true = {1.0, -2.0, 0.5, 3.0}, b(4, 0.0), x(4, 0.0);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) b[i] += a[i][j] * x_true[j];
  LU lu(a);
  lu.solve(b, x);
  double err = 0.0;
  for (int i = 0; i < 4; i++) err = std::fmax(err, std::fabs(x[i] - x_true[i]));
  emit("lu_solve", err < 1e-13, err);
  MatrixDoub inv;
  lu.inverse(inv);
  double ie = 0.0;
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      double s = 0.0;
      for (int k = 0; k < 4; k++) s += a[i][k] * inv[k][j];
      ie = std::fmax(ie, std::fabs(s - (i == j ? 1.0 : 0.0)));
    }
  emit("lu_inverse", ie < 1e-13, ie);
  emit("lu_det", true, lu.det());
  VecDoub xp = x;
  xp[0] += 1e-9;
  lu.mprove(b, xp);
  double me = 0.0;
  for (int i = 0; i < 4; i++) me = std::fmax(me, std::fabs(xp[i] - x_true[i]));
  emit("lu_mprove", me < 1e-13, me);
  // a zero row: the reference exits, the class throws
  bool threw = false;
  try {
    MatrixDoub z = {{1, 2}, {0, 0}};
    LU bad(z);
  } catch (const mts::Error &e) {
    threw = e.code() == MTS_E_ARG && std::string(e.what()).find("Singular") != std::string::npos;
  }
  emit("lu_singular_throws", threw, 0.0);
  // Mesh and Intersections on the host
  Surface_Ellipsoid ell({0.7, 0.5, 0.4, 1 - std::ldexp(1.0, -20)});
  Mesh mesh({-1, 1, -1, 1, -1, 1}, {33, 33, 33}, ell);
  int inside = 0;
  for (int i = 0; i < 33 * 33 * 33; i++) inside += mesh.mesh_value[i] < 0;
  emit("mesh_inside_nodes", inside > 0 && mesh.dx == 2.0 / 32, inside);
  Beta_SinCos beta(10.0, 1.0);
  std::ofstream sink("/dev/null");
  Intersections inter(ell, mesh, beta, 2, 2, sink);
  size_t nc = 0;
  for (auto &row : inter.ifpx)
    for (auto &line : row) nc += line.size();
  emit("intersections_x", nc > 0, (double)nc);
  // interface through a node: the reference exits, the class throws MTS_E_GEOMETRY
  threw = false;
  try {
    Surface_Ellipsoid on_node({1, 1, 1, 0.5});
    Mesh m2({-1, 1, -1, 1, -1, 1}, {33, 33, 33}, on_node);
  } catch (const mts::Error &e) {
    threw = e.code() == MTS_E_GEOMETRY;
  }
  emit("mesh_on_node_throws", threw, 0.0);
  return 0;
}

// max |a - b| over all nodes
static double field_diff(const CubicDoub &a, const CubicDoub &b) {
  double d = 0.0;
  for (size_t i = 0; i < a.size(); i++)
    for (size_t j = 0; j < a[i].size(); j++)
      for (size_t k = 0; k < a[i][j].size(); k++) d = std::fmax(d, std::fabs(a[i][j][k] - b[i][j][k]));
  return d;
}

// The same solve through the C ABI directly, in [ix][iy][iz] order.
static CubicDoub c_abi_run(Intersections &inter, Mesh &mesh, int method, const VecDoub &time,
                           int steps) {
  mts_ctx *ctx = nullptr;
  mts_instance *in = nullptr;
  check(mts_ctx_create(0, &ctx));
  check(mts_instance_create(ctx, inter.handle(), &inter.physics(), method, time[2], &in));
  check(mts_initialize(in, 0.0));
  double t = 0.0;
  for (int j = 0; j < steps; j++) {
    t += time[2];
    check(mts_step(in, t));
  }
  std::vector<double> flat((size_t)mesh.nx * mesh.ny * mesh.nz);
  check(mts_field_download(in, flat.data(), MTS_LAYOUT_REF));
  mts_instance_destroy(in);
  mts_ctx_destroy(ctx);
  CubicDoub u(mesh.nx, MatrixDoub(mesh.ny, VecDoub(mesh.nz)));
  for (int i = 0; i < mesh.nx; i++)
    for (int j = 0; j < mesh.ny; j++)
      for (int k = 0; k < mesh.nz; k++) u[i][j][k] = flat[((size_t)i * mesh.ny + j) * mesh.nz + k];
  return u;
}

static int gpu_checks() {
  Surface_Ellipsoid ell({0.7, 0.5, 0.4, 1 - std::ldexp(1.0, -20)});
  Mesh mesh({-1, 1, -1, 1, -1, 1}, {41, 41, 41}, ell);
  Beta_SinCos beta(10.0, 1.0);
  std::ofstream sink("/dev/null");
  Intersections inter(ell, mesh, beta, 2, 2, sink);
  const VecDoub time = {0.0, 0.25, 0.05};
  const int steps = 5;
  // LOD-CN with the host copy updated every step, and with the field kept on
  // the device: the same arithmetic, so the same field
  CubicDoub u1, u2;
  {
    LOD lod(inter, mesh, beta, time);
    Equation e0(0.0, beta);
    lod.Initialization(e0, u1);
    char m = 'C';
    double t = 0.0;
    for (int j = 0; j < steps; j++) {
      t += time[2];
      Equation e(t, beta);
      lod.Solve_2nd(e, inter, u1, beta, m);
    }
  }
  {
    LOD lod(inter, mesh, beta, time);
    lod.set_host_sync(false);
    Equation e0(0.0, beta);
    lod.Initialization(e0, u2);
    char m = 'C';
    double t = 0.0;
    for (int j = 0; j < steps; j++) {
      t += time[2];
      Equation e(t, beta);
      lod.Solve_2nd(e, inter, u2, beta, m);
    }
    lod.sync_host(u2);
  }
  const double d = field_diff(u1, u2);
  emit("lod_host_sync_same_field", d == 0.0, d);
  // the classes add no arithmetic: the same field as the C ABI on the same inputs
  const double dc = field_diff(u1, c_abi_run(inter, mesh, MTS_LOD_CN, time, steps));
  emit("lod_same_as_c_abi", dc == 0.0, dc);
  // TS through the three-equation signature
  {
    CubicDoub u;
    TS ts(inter, mesh, beta, time);
    Equation e0(0.0, beta);
    ts.Initialization(e0, u);
    double tn = 0.0;
    for (int j = 0; j < steps; j++) {
      const double tp = tn + time[2];
      Equation a(tn, beta), h(tn + 0.5 * time[2], beta), c(tp, beta);
      ts.Solve_2nd(a, h, c, inter, u, beta);
      tn = tp;
    }
    const double e = field_diff(u, c_abi_run(inter, mesh, MTS_TS, time, steps));
    emit("ts_same_as_c_abi", e == 0.0, e);
  }
  // Douglas-ADI on an interface with corner pairs is switched off: the
  // constructor throws MTS_E_DISABLED and runs no kernel
  {
    bool threw = false;
    Surface_Ellipsoid e2({0.7, 0.5, 0.4, 1 - std::ldexp(1.0, -20)});
    Mesh m2({-1, 1, -1, 1, -1, 1}, {65, 65, 65}, e2);
    Intersections i2(e2, m2, beta, 2, 1, sink);
    try {
      Douglas_ADI adi(i2, m2, beta, {0.0, 0.1, 1.0 / 32}, 2);
    } catch (const mts::Error &e) {
      threw = e.code() == MTS_E_DISABLED;
    }
    emit("adi_corners_refused", threw, 0.0);
  }
  return 0;
}

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "host";
  try {
    if (mode == "host") return host_checks();
    if (mode == "gpu") return gpu_checks();
  } catch (const std::exception &e) {
    std::printf("{\"check\": \"exception\", \"ok\": false, \"message\": \"%s\"}\n", e.what());
    return 1;
  }
  std::fprintf(stderr, "usage: %s host|gpu\n", argv[0]);
  return 2;
}
