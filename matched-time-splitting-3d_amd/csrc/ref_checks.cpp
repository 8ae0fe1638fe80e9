// Where the reference would have stopped a run that this build completes.
//
// For every crossing the reference's set-up prepares numeric tangential
// (eta, tau) flux-jump approximations (Auxiliary_eta_* / Auxiliary_tau_*,
// spatial/intersection.cc:2746-3660, called from Getdata_* :346-779), and
// every step's Refresh_Jump evaluates them (Eta_* / Tau_*, :3670-4438).
// Their results reach no output in the reference's approximate-jump mode
// (constant.h:16, JP = 'a'), so this build computes none of them; but their
// geometric checks call exit(0), and a caller of this build should be able to
// learn that the reference would not have completed.  This file restates
// exactly those checks: the two auxiliary points of each tangent line must
// lie inside the domain along the line's axis, each of the two neighbouring
// grid lines they lie on must hold three consecutive nodes of one side
// (Search_indx :2613-2730), and the tangent frame component the
// approximation's orientation is read from must not be zero.
#include <cmath>
#include <cstdio>
#include <string>

#include "host_setup.hpp"
#include "ref_checks.hpp"

namespace mts {

namespace {

// Search_indx (intersection.cc:2613-2730): whether the grid line along
// `axis` through (ix, iy, iz) (the index along the axis ignored) holds three
// consecutive nodes outside, and three inside.  The reference scans start
// indices i < nx - 2 for both x and y lines (:2634, :2661) and i < nz - 2 for
// z lines; a y scan is bounded by the y line here as well.
void search_runs(const SetupResult &s, int axis, int ix, int iy, int iz, bool &out3, bool &in3) {
  const Grid &g = s.grid;
  out3 = in3 = false;
  const int n = axis == 2 ? g.n[2] : (axis == 1 ? std::min(g.n[0], g.n[1]) : g.n[0]);
  for (int i = 0; i + 2 < n; i++) {
    int o = 0, m = 0;
    for (int t = 0; t < 3; t++) {
      const int8_t v = axis == 0 ? s.ind[g.idx(i + t, iy, iz)]
                                 : axis == 1 ? s.ind[g.idx(ix, i + t, iz)]
                                             : s.ind[g.idx(ix, iy, i + t)];
      o += v > 0;
      m += v < 0;
    }
    if (o == 3) out3 = true;
    if (m == 3) in3 = true;
  }
}

const char *const kAxis[3] = {"x", "y", "z"};

}  // namespace

// One auxiliary-point set-up (eta or tau) of crossing c along direction d.
// Returns the stop kind or -1.
static int aux_check(const SetupResult &s, const Crossing &c, bool eta, std::string &msg) {
  const Grid &g = s.grid;
  const int d = c.dir;
  const double *a = c.normal;
  const double x0 = c.coord[0], y0 = c.coord[1], z0 = c.coord[2];
  // the line's node indices; the other axis the tangent line steps along
  int ix = 0, iy = 0, iz = 0;
  if (d == 0) { iy = c.j; iz = c.k; }
  else if (d == 1) { ix = c.j; iz = c.k; }
  else { ix = c.j; iy = c.k; }
  // auxiliary points on the neighbouring lines (upper +1, lower -1) of the
  // stepping axis, as coordinates along d (intersection.cc:2804, 3007, 3149,
  // 3291, 3433, 3575)
  int step;  // the stepping axis
  double au, al;
  if (d == 0 && eta) {
    step = 2;
    au = (a[0] * x0 + a[2] * (z0 - g.c[2][iz + 1])) / a[0];
    al = (a[0] * x0 + a[2] * (z0 - g.c[2][iz - 1])) / a[0];
  } else if (d == 0) {
    step = 1;
    au = (a[0] * x0 + a[1] * (y0 - g.c[1][iy + 1])) / a[0];
    al = (a[0] * x0 + a[1] * (y0 - g.c[1][iy - 1])) / a[0];
  } else if (d == 1 && eta) {
    step = 2;
    au = (a[1] * y0 + a[2] * (z0 - g.c[2][iz + 1])) / a[1];
    al = (a[1] * y0 + a[2] * (z0 - g.c[2][iz - 1])) / a[1];
  } else if (d == 1) {
    step = 0;
    au = (a[1] * y0 + a[0] * (x0 - g.c[0][ix + 1])) / a[1];
    al = (a[1] * y0 + a[0] * (x0 - g.c[0][ix - 1])) / a[1];
  } else if (eta) {
    step = 1;
    au = (a[2] * z0 + a[1] * (y0 - g.c[1][iy + 1])) / a[2];
    al = (a[2] * z0 + a[1] * (y0 - g.c[1][iy - 1])) / a[2];
  } else {
    step = 0;
    au = (a[2] * z0 + a[0] * (x0 - g.c[0][ix + 1])) / a[2];
    al = (a[2] * z0 + a[0] * (x0 - g.c[0][ix - 1])) / a[2];
  }
  const std::vector<double> &cd = g.c[d];
  const int n = g.n[d];
  const char *what = eta ? "eta" : "tau";
  static const char *const kPlane[3] = {"Plane X", "Plane Y", "Plane Z"};
  // the reference's range test (strict comparisons: NaN passes)
  if (au < cd[0] || au > cd[n - 1] || al < cd[0] || al > cd[n - 1]) {
    char b[160];
    std::snprintf(b, sizeof b,
                  "Auxilary nodes in %s direction are out of boundary on %s direction of %s", what,
                  kAxis[d], kPlane[3 - d - step]);
    msg = b;
    return kStopAuxOutside;
  }
  // its index search (xi[i] < p < xi[i + 1]) leaves the index unset when a
  // point sits on a node or is NaN: undefined in the reference
  auto inside_cell = [&](double p) {
    for (int i = 0; i + 1 < n; i++)
      if (cd[i] < p && p < cd[i + 1]) return true;
    return false;
  };
  if (!inside_cell(au) || !inside_cell(al)) {
    msg = std::string("auxiliary point in ") + what + " direction on a grid node or undefined "
          "(the reference's node search leaves its index unset)";
    return kStopAuxOnNode;
  }
  // Search_indx on the upper and lower neighbouring lines
  int ux = ix, uy = iy, uz = iz, lx = ix, ly = iy, lz = iz;
  if (step == 0) { ux = ix + 1; lx = ix - 1; }
  if (step == 1) { uy = iy + 1; ly = iy - 1; }
  if (step == 2) { uz = iz + 1; lz = iz - 1; }
  bool uo, ui, lo, li;
  search_runs(s, d, ux, uy, uz, uo, ui);
  search_runs(s, d, lx, ly, lz, lo, li);
  if (!(ui && li) && !(uo && lo)) {
    msg = std::string(eta ? "Eta" : "Tau") + " approximation fails";
    return kStopNoSameSide;
  }
  return -1;
}

// The frame component Eta_* / Tau_* orient their approximation by
// (intersection.cc:3712, 3846, 3975, 4104, 4233, 4365).
static int frame_check(const Crossing &c, bool eta, std::string &msg) {
  // [eta][d]: Tau_x p[2][1], Tau_y p[2][0], Tau_z p[2][0]; Eta_x p[1][2],
  // Eta_y p[1][2], Eta_z p[1][1]
  static const int kRow[2][3] = {{2, 2, 2}, {1, 1, 1}};
  static const int kCol[2][3] = {{1, 0, 0}, {2, 2, 1}};
  const double v = c.p[kRow[eta][c.dir]][kCol[eta][c.dir]];
  if (v > 0 || v < 0) return -1;
  msg = std::string(kAxis[c.dir]) + "0 is on " + kAxis[c.dir] + "-axis";
  return kStopFrameZero;
}

void reference_stops(const SetupResult &s, RefStops &out) {
  out = RefStops();
  const Grid &g = s.grid;
  // lines in the reference's loop order: z lines (ix outer, iy inner), then y
  // lines (ix, iz), then x lines (iy, iz); a line's crossings in order, a
  // crossing's eta before its tau (Setup_Intersections :79-291)
  auto for_lines = [&](auto &&fn) {
    const int dirs[3] = {2, 1, 0};
    for (int d : dirs) {
      const int no = d == 0 ? g.n[1] : g.n[0];  // outer index range
      const int ni = d == 2 ? g.n[1] : g.n[2];  // inner
      for (int o = 1; o < no - 1; o++)
        for (int i = 1; i < ni - 1; i++) {
          // the line's (j, k): x lines (iy, iz), y lines (ix, iz), z lines (ix, iy)
          const int lid = s.line_id(d, o, i);
          for (int ci = s.line_ptr[d][lid]; ci < s.line_ptr[d][lid + 1]; ci++) fn(d, ci);
        }
    }
  };
  auto note = [&](int kind, int d, int ci, const std::string &msg) {
    out.count[kind]++;
    if (out.first_kind < 0 && kind != kStopAuxOnNode) {  // (not an exit)
      out.first_kind = kind;
      out.first_dir = d;
      out.first_index = ci;
      out.message = msg;
    }
  };
  // set-up (Getdata_*): both auxiliary set-ups of every crossing
  for_lines([&](int d, int ci) {
    const Crossing &c = s.cross[d][ci];
    for (int e = 1; e >= 0; e--) {
      std::string msg;
      const int k = aux_check(s, c, e == 1, msg);
      if (k >= 0) note(k, d, ci, msg);
    }
  });
  // first step (Refresh_Jump): Eta_*, Tau_* per crossing
  for_lines([&](int d, int ci) {
    const Crossing &c = s.cross[d][ci];
    for (int e = 1; e >= 0; e--) {
      std::string msg;
      const int k = frame_check(c, e == 1, msg);
      if (k >= 0) note(k, d, ci, msg);
    }
  });
}

}  // namespace mts
