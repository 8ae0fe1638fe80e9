// End-of-run interface diagnostics on the host (FP64): the fictitious-point
// and jump errors of the reference's report (Refresh_Fp / Error_Fp,
// spatial/intersection.cc:1841-2281 and 4667-4771; Refresh_Jump's jump.err,
// intersection.cc:2294-2490, and Error_Jump, :4594-4665), for the
// manufactured solution the kernels evaluate.
#include "diagnostics.hpp"

#include <cmath>

namespace mts {
namespace {

// u- = e sin x sin y sin z, u+ = e cos x cos y cos z and their jumps, in the
// operation order of the checker's plugin of the same solution.
struct Exact {
  double e;
  const BetaModel &beta;
  double outer(double x, double y, double z) const { return e * std::cos(x) * std::cos(y) * std::cos(z); }
  double inner(double x, double y, double z) const { return e * std::sin(x) * std::sin(y) * std::sin(z); }
  double jump_u(double x, double y, double z) const { return outer(x, y, z) - inner(x, y, z); }
  // [beta du/dx_d]
  double jump_bd(int d, double x, double y, double z) const {
    double og, ig;
    if (d == 0) {
      og = -e * std::sin(x) * std::cos(y) * std::cos(z);
      ig = e * std::cos(x) * std::sin(y) * std::sin(z);
    } else if (d == 1) {
      og = -e * std::cos(x) * std::sin(y) * std::cos(z);
      ig = e * std::sin(x) * std::cos(y) * std::sin(z);
    } else {
      og = -e * std::cos(x) * std::cos(y) * std::sin(z);
      ig = e * std::sin(x) * std::sin(y) * std::cos(z);
    }
    return beta.outside(x, y, z) * og - beta.inside(x, y, z) * ig;
  }
};

void line_axes(int d, int &ja, int &ka) {
  if (d == 0) { ja = 1; ka = 2; }
  else if (d == 1) { ja = 0; ka = 2; }
  else { ja = 0; ka = 1; }
}

// Node i of line (d, j, k): its exact value on its own side (u) and on the
// other side (invu), as Refresh_Fp tabulates them.
struct LineNodes {
  const SetupResult &s;
  const Exact &ex;
  int d, j, k;
  void xyz(int i, double &x, double &y, double &z, int &ix, int &iy, int &iz) const {
    int idx[3];
    int ja, ka;
    line_axes(d, ja, ka);
    idx[d] = i;
    idx[ja] = j;
    idx[ka] = k;
    ix = idx[0], iy = idx[1], iz = idx[2];
    x = s.grid.c[0][ix], y = s.grid.c[1][iy], z = s.grid.c[2][iz];
  }
  double u(int i) const {
    double x, y, z;
    int ix, iy, iz;
    xyz(i, x, y, z, ix, iy, iz);
    return s.ind[s.grid.idx(ix, iy, iz)] > 0 ? ex.outer(x, y, z) : ex.inner(x, y, z);
  }
  double invu(int i) const {
    double x, y, z;
    int ix, iy, iz;
    xyz(i, x, y, z, ix, iy, iz);
    return s.ind[s.grid.idx(ix, iy, iz)] > 0 ? ex.inner(x, y, z) : ex.outer(x, y, z);
  }
};

double rel_err(double ref, double v) { return std::fabs((ref - v) / ref); }

}  // namespace

void crossing_errors(const SetupResult &s, const BetaModel &beta, int d, double t,
                     std::vector<double> &errl, std::vector<double> &errr,
                     std::vector<double> &jerr) {
  const Exact ex{std::exp(-t), beta};
  const std::vector<Crossing> &cr = s.cross[d];
  errl.assign(cr.size(), 0.0);
  errr.assign(cr.size(), 0.0);
  jerr.assign(cr.size(), 0.0);
  for (size_t c = 0; c < cr.size(); c++) {
    const Crossing &x = cr[c];
    // jump error of the GFM u_dir = P^{-1}[d][0] (P row 0 . [beta grad u])
    const double bxi = x.p[0][0] * ex.jump_bd(0, x.coord[0], x.coord[1], x.coord[2]) +
                       x.p[0][1] * ex.jump_bd(1, x.coord[0], x.coord[1], x.coord[2]) +
                       x.p[0][2] * ex.jump_bd(2, x.coord[0], x.coord[1], x.coord[2]);
    const double udir = x.pinv_d0 * bxi;
    const double bd = ex.jump_bd(d, x.coord[0], x.coord[1], x.coord[2]);
    jerr[c] = std::fabs(udir - bd) / bd;
  }
  for (size_t c = 0; c < cr.size();) {
    const Crossing &L = cr[c];
    const LineNodes ln{s, ex, d, L.j, L.k};
    const int ll = L.left_loc;
    if (L.id > 0) {
      // irregular: 4 grid weights on nodes ll-1..ll+2, then [u], [beta u_d]
      const double ju = ex.jump_u(L.coord[0], L.coord[1], L.coord[2]);
      const double jb = ex.jump_bd(d, L.coord[0], L.coord[1], L.coord[2]);
      double sl = 0.0, sr = 0.0;
      sl += L.wl[4] * ju;
      sl += L.wl[5] * jb;
      sr += L.wr[4] * ju;
      sr += L.wr[5] * jb;
      for (int q = 0; q < 4; q++) {
        const double v = ln.u(ll - 1 + q);
        sl += L.wl[q] * v;
        sr += L.wr[q] * v;
      }
      errl[c] = rel_err(ln.invu(ll), sl);
      errr[c] = rel_err(ln.invu(ll + 1), sr);
      c += 1;
    } else {
      // corner pair L, R: 5 grid weights on nodes ll-1..ll+3, then the jumps
      // of L and of R; the middle fictitious point is shared
      const Crossing &R = cr[c + 1];
      const double juL = ex.jump_u(L.coord[0], L.coord[1], L.coord[2]);
      const double jbL = ex.jump_bd(d, L.coord[0], L.coord[1], L.coord[2]);
      const double juR = ex.jump_u(R.coord[0], R.coord[1], R.coord[2]);
      const double jbR = ex.jump_bd(d, R.coord[0], R.coord[1], R.coord[2]);
      double lsl = 0.0, lsr = 0.0, rsr = 0.0;
      lsl += L.wl[5] * juL;
      lsl += L.wl[6] * jbL;
      lsl += L.wl[7] * juR;
      lsl += L.wl[8] * jbR;
      lsr += L.wr[5] * juL;
      lsr += L.wr[6] * jbL;
      lsr += L.wr[7] * juR;
      lsr += L.wr[8] * jbR;
      rsr += R.wr[5] * juL;
      rsr += R.wr[6] * jbL;
      rsr += R.wr[7] * juR;
      rsr += R.wr[8] * jbR;
      for (int q = 0; q < 5; q++) {
        const double v = ln.u(ll - 1 + q);
        lsl += L.wl[q] * v;
        lsr += L.wr[q] * v;
        rsr += R.wr[q] * v;
      }
      errl[c] = rel_err(ln.invu(ll), lsl);
      errr[c] = rel_err(ln.invu(ll + 1), lsr);
      errl[c + 1] = errr[c];
      errr[c + 1] = rel_err(ln.invu(ll + 2), rsr);
      c += 2;
    }
  }
}

void interface_diagnostics(const SetupResult &s, const BetaModel &beta, double t, double fp[2],
                           double jump[2]) {
  std::vector<double> el[3], er[3], je[3];
  for (int d = 0; d < 3; d++) crossing_errors(s, beta, d, t, el[d], er[d], je[d]);
  double fmax = 0.0, fsum = 0.0, jmax = 0.0, jsum = 0.0;
  long fno = 0, jno = 0;
  // the reference's loop order: z lines, y lines, x lines, each with the
  // first fixed index outer
  for (int d : {2, 1, 0}) {
    int ja, ka;
    line_axes(d, ja, ka);
    const int nj = s.grid.n[ja], nk = s.grid.n[ka];
    for (int j = 1; j < nj - 1; j++)
      for (int k = 1; k < nk - 1; k++) {
        const int lid = s.line_id(d, j, k);
        for (int c = s.line_ptr[d][lid]; c < s.line_ptr[d][lid + 1]; c++) {
          fno += 2;
          fsum += el[d][c] * el[d][c];
          fsum += er[d][c] * er[d][c];
          if (std::fabs(el[d][c]) > fmax) fmax = std::fabs(el[d][c]);
          if (std::fabs(er[d][c]) > fmax) fmax = std::fabs(er[d][c]);
          jno += 1;
          jsum += je[d][c] * je[d][c];
          if (std::fabs(je[d][c]) > jmax) jmax = std::fabs(je[d][c]);
        }
      }
  }
  fp[0] = fmax;
  fp[1] = std::sqrt(fsum / fno);
  jump[0] = jmax;
  jump[1] = std::sqrt(jsum / jno);
}

}  // namespace mts
