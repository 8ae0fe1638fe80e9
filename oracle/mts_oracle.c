/*
 * ORACLE (test infrastructure only; see mts_oracle.h).
 *
 * CPU restatement of the reference's per-time-step path in plain C99.  Every
 * routine names the reference code it restates (paths relative to the
 * reference's src/).  The loops follow the reference's structure: full-field
 * right-hand-side passes, fictitious-point patches per crossing, per-line
 * tridiagonal assembly with Convert2Tri row reduction and a sequential Thomas
 * solve.  The arithmetic order inside each expression follows the reference.
 *
 * The manufactured solution and the diffusion coefficients are the ones the
 * benchmark configurations name (u- = e^{-t} sin x sin y sin z,
 * u+ = e^{-t} cos x cos y cos z; beta constant per side or
 * beta- = c (1 + 0.5 sin x cos y)); the reference has no plugin for them, so
 * oracle/ref_harness.cc adds the same formulas to the reference build.
 *
 * The one deliberate difference: the directional flux jump is computed in
 * closed form, u_dir = n_d (n . [beta grad u]), where the reference inverts
 * its 3x3 local frame P with LU and takes P^{-1}[d][0]
 * (intersection.cc:2345-2352).  The first column of P^{-1} is n for any frame
 * whose first row is the unit normal, so the two agree to rounding.
 */
#include "mts_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int j, k, id, ll;
  double coord[3], normal[3], wl[9], wr[9];
  double ju, jb; /* refreshed each step */
} Cr;

typedef struct {
  int n[3];
  double h[3];
  double *c[3];
  const signed char *ind;
  int bk;
  double b0, b1;
  Cr *cr[3];
  int ncr[3];
  int *lptr[3];
  int nj[3], nk[3];
  size_t N;
  /* line work arrays */
  double *la, *lb, *lc, *lr, *lu, *gam;
} Prob;

static size_t IDX(const Prob *P, int ix, int iy, int iz) {
  return ((size_t)iz * P->n[1] + iy) * P->n[0] + ix;
}

/* axes of a d-line: fixed indices (j along ja, k along ka) */
static void axes(int d, int *ja, int *ka) {
  if (d == 0) { *ja = 1; *ka = 2; }
  else if (d == 1) { *ja = 0; *ka = 2; }
  else { *ja = 0; *ka = 1; }
}

static void node_of(int d, int j, int k, int i, int q[3]) {
  int ja, ka;
  axes(d, &ja, &ka);
  q[ja] = j;
  q[ka] = k;
  q[d] = i;
}

static size_t NODE(const Prob *P, int d, int j, int k, int i) {
  int q[3];
  node_of(d, j, k, i, q);
  return IDX(P, q[0], q[1], q[2]);
}

/* ---------------------------------------------------------------- physics */
static double beta_in(const Prob *P, double x, double y, double z) {
  (void)z;
  if (P->bk == 1) return P->b0 * (1.0 + 0.5 * sin(x) * cos(y));
  return P->b0;
}
static double beta_out(const Prob *P, double x, double y, double z) {
  (void)x; (void)y; (void)z;
  return P->b1;
}
static double u_in(double e, double x, double y, double z) { return e * sin(x) * sin(y) * sin(z); }
static double u_out(double e, double x, double y, double z) { return e * cos(x) * cos(y) * cos(z); }

static double f_in(const Prob *P, double e, double x, double y, double z) {
  double v = (3.0 * beta_in(P, x, y, z) - 1.0) * u_in(e, x, y, z);
  if (P->bk == 1) {
    double bx = 0.5 * P->b0 * cos(x) * cos(y);
    double by = -0.5 * P->b0 * sin(x) * sin(y);
    v = v - bx * (e * cos(x) * sin(y) * sin(z)) - by * (e * sin(x) * cos(y) * sin(z));
  }
  return v;
}
static double f_out(const Prob *P, double e, double x, double y, double z) {
  return (3.0 * beta_out(P, x, y, z) - 1.0) * u_out(e, x, y, z);
}

/* Operator_weights_{x,y,z} (douglas_adi.cc:1444-1539): centre node's side */
static void opw(const Prob *P, int d, int ix, int iy, int iz, double vd[3]) {
  double p1[3] = {P->c[0][ix], P->c[1][iy], P->c[2][iz]};
  double p2[3] = {P->c[0][ix], P->c[1][iy], P->c[2][iz]};
  double b1, b2, h = P->h[d];
  p1[d] = p1[d] - h / 2;
  p2[d] = p2[d] + h / 2;
  if (P->ind[IDX(P, ix, iy, iz)] > 0) {
    b1 = beta_out(P, p1[0], p1[1], p1[2]);
    b2 = beta_out(P, p2[0], p2[1], p2[2]);
  } else {
    b1 = beta_in(P, p1[0], p1[1], p1[2]);
    b2 = beta_in(P, p2[0], p2[1], p2[2]);
  }
  vd[0] = b1 / (h * h);
  vd[1] = -(b1 + b2) / (h * h);
  vd[2] = b2 / (h * h);
}

static void opw_row(const Prob *P, int d, int j, int k, int i, double vd[3]) {
  int q[3];
  node_of(d, j, k, i, q);
  opw(P, d, q[0], q[1], q[2], vd);
}

/* Refresh_Jump live part (intersection.cc:2294-2490): [u] and u_dir */
static void refresh_jumps(Prob *P, double e) {
  for (int d = 0; d < 3; d++)
    for (int i = 0; i < P->ncr[d]; i++) {
      Cr *c = &P->cr[d][i];
      double x = c->coord[0], y = c->coord[1], z = c->coord[2];
      double bi = beta_in(P, x, y, z), bo = beta_out(P, x, y, z);
      c->ju = u_out(e, x, y, z) - u_in(e, x, y, z);
      double jbx = bo * (-e * sin(x) * cos(y) * cos(z)) - bi * (e * cos(x) * sin(y) * sin(z));
      double jby = bo * (-e * cos(x) * sin(y) * cos(z)) - bi * (e * sin(x) * cos(y) * sin(z));
      double jbz = bo * (-e * cos(x) * cos(y) * sin(z)) - bi * (e * sin(x) * sin(y) * cos(z));
      double xi = c->normal[0] * jbx + c->normal[1] * jby + c->normal[2] * jbz;
      c->jb = c->normal[d] * xi;
    }
}

/* corner jump combination, e.g. douglas_adi.cc:1218-1221 */
static double JL(const double *w, const Cr *L, const Cr *R) {
  return L->ju * w[5] + L->jb * w[6] + R->ju * w[7] + R->jb * w[8];
}

/* ---------------------------------------------------------------- TDMA */
int mts_oracle_tdma(int n, const double *a, const double *b, const double *c, const double *r,
                    double *u) {
  double *gam = (double *)malloc(sizeof(double) * (size_t)n);
  double bet;
  if (!gam) return -1;
  if (b[0] == 0.0) {
    free(gam);
    return -1;
  }
  u[0] = r[0] / (bet = b[0]);
  for (int j = 1; j < n; j++) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    if (bet == 0.0) {
      free(gam);
      return -1;
    }
    u[j] = (r[j] - a[j] * u[j - 1]) / bet;
  }
  for (int j = n - 2; j >= 0; j--) u[j] -= gam[j + 1] * u[j + 1];
  free(gam);
  return 0;
}

/* ------------------------------------------------ alternative line solvers
 * Not part of the reference.  They let the tests separate rounding from
 * error: the same step is solved with a different elimination and the
 * differences measured against the conditioning of the line systems.
 *
 * gtsv: Gaussian elimination with partial (row-interchange) pivoting for a
 * tridiagonal matrix, the algorithm of LAPACK dgtsv. */
static int gtsv(int n, const double *a, const double *b, const double *c, const double *r,
                double *u) {
  double *dl = (double *)malloc(sizeof(double) * (size_t)n * 4);
  if (!dl) return -1;
  double *dd = dl + n, *du = dd + n, *du2 = du + n;
  for (int i = 0; i < n; i++) {
    dl[i] = i + 1 < n ? a[i + 1] : 0.0; /* sub-diagonal entry of row i+1 */
    dd[i] = b[i];
    du[i] = c[i];
    du2[i] = 0.0;
    u[i] = r[i];
  }
  for (int i = 0; i < n - 1; i++) {
    if (fabs(dd[i]) >= fabs(dl[i])) {
      if (dd[i] == 0.0) {
        free(dl);
        return -1;
      }
      double f = dl[i] / dd[i];
      dd[i + 1] -= f * du[i];
      u[i + 1] -= f * u[i];
      dl[i] = 0.0;
    } else {
      double f = dd[i] / dl[i];
      dd[i] = dl[i];
      double tmp = dd[i + 1];
      dd[i + 1] = du[i] - f * tmp;
      if (i < n - 2) {
        dl[i] = du[i + 1];
        du[i + 1] = -f * dl[i];
      }
      du2[i] = i < n - 2 ? dl[i] : 0.0;
      du[i] = tmp;
      tmp = u[i];
      u[i] = u[i + 1];
      u[i + 1] = tmp - f * u[i + 1];
    }
  }
  if (dd[n - 1] == 0.0) {
    free(dl);
    return -1;
  }
  u[n - 1] /= dd[n - 1];
  if (n > 1) u[n - 2] = (u[n - 2] - du[n - 2] * u[n - 1]) / dd[n - 2];
  for (int i = n - 3; i >= 0; i--) u[i] = (u[i] - du[i] * u[i + 1] - du2[i] * u[i + 2]) / dd[i];
  free(dl);
  return 0;
}

/* Model of the GPU's line solve, operation for operation, with fused
 * multiply-adds where the GPU fuses (the GPU forms reciprocals with a refined
 * hardware estimate where this model divides).
 * A line with a row that is not diagonally dominant (|b| < |a| + |c|) is
 * solved with the Thomas algorithm.  Any other line: the boundary rows 0 and
 * n-1 (identity rows) are moved to the right-hand side; the interior rows
 * 1..n-2 are split into p segments of m rows (the last one mlast rows), one
 * per lane of a group of L lanes (split_line: the smallest of 16, 32, 64
 * lanes and 4, 8, 16 rows per lane that fits, with no one-row segment); each
 * segment is eliminated down and up keeping its first unknown as a
 * parameter, folded to one equation per lane in the segment-end unknowns,
 * solved by parallel cyclic reduction over the L lanes (idle lanes are
 * identity rows), then back-substituted.  When every segment but the last
 * has exactly the register bound mmax of rows, the last segment is padded to
 * mmax rows with identity rows (zero right-hand side), as the kernel's
 * unguarded solve eliminates them. */
typedef struct {
  int L, m, mlast, p, mmax;
} Split;

static int split_line(int n, Split *q) {
  static const int kL[3] = {16, 32, 64}, kM[3] = {4, 8, 16};
  const int ni = n - 2;
  if (ni < 2 || n > 1024) return -1;
  for (int li = 0; li < 3; li++) {
    const int L = kL[li];
    int m = (ni + L - 1) / L;
    if (m < 2) m = 2;
    int p = (ni + m - 1) / m;
    int mlast = ni - (p - 1) * m;
    if (mlast == 1) {
      p -= 1;
      mlast = m + 1;
    }
    if (p > L) continue;
    const int mx = m > mlast ? m : mlast;
    for (int mi = 0; mi < 3; mi++)
      if (mx <= kM[mi]) {
        q->L = L;
        q->m = m;
        q->mlast = mlast;
        q->p = p;
        q->mmax = kM[mi];
        return 0;
      }
  }
  return -1;
}

static int part_solve(int n, const double *A, const double *B, const double *C, const double *R,
                      double *U) {
  enum { LMAX = 64 };
  Split q;
  for (int i = 0; i < n; i++)
    if (fabs(B[i]) < fabs(A[i]) + fabs(C[i])) return mts_oracle_tdma(n, A, B, C, R, U);
  if (split_line(n, &q) != 0) return -1;
  /* line systems: rows 0 and n-1 carry the boundary values */
  if (A[0] != 0.0 || B[0] != 1.0 || C[0] != 0.0 || A[n - 1] != 0.0 || B[n - 1] != 1.0 ||
      C[n - 1] != 0.0)
    return -1;
  /* full split (every segment but the last has exactly mmax rows): the last
   * segment is eliminated over m rows, its rows past the line's last interior
   * row being identity rows with a zero right-hand side, as the kernel's
   * unguarded solve does */
  const int full = q.m == q.mmax;
  const int np = n + q.m;
  double *w = (double *)malloc(sizeof(double) * (size_t)np * 4);
  if (!w) return -1;
  double *a = w, *c = w + np, *d = w + 2 * np, *b = w + 3 * np;
  double A1[LMAX], B1[LMAX], C1[LMAX], D1[LMAX], A2[LMAX], C2[LMAX], D2[LMAX];
  double ea[LMAX], eb[LMAX], ec[LMAX], ed[LMAX], xe[LMAX];
  const int L = q.L;
  for (int i = 0; i < np; i++) {
    a[i] = i < n ? A[i] : 0.0;
    b[i] = i < n ? B[i] : 1.0;
    c[i] = i < n ? C[i] : 0.0;
    d[i] = i < n - 1 ? R[i] : 0.0;
  }
  /* boundary rows are identity rows: their values are known */
  const double x0 = R[0], xn = R[n - 1];
  d[1] = d[1] - a[1] * x0;
  a[1] = 0.0;
  d[n - 2] = d[n - 2] - c[n - 2] * xn;
  c[n - 2] = 0.0;
  for (int k = 0; k < L; k++) {
    A1[k] = 0.0; B1[k] = 1.0; C1[k] = 0.0; D1[k] = 0.0;
    A2[k] = 0.0; C2[k] = 0.0; D2[k] = 0.0;
    if (k >= q.p) continue;
    const int s = 1 + k * q.m, m = k == q.p - 1 && !full ? q.mlast : q.m;
    double *ra = a + s, *rc = c + s, *rd = d + s;
    const double *rb = b + s;
    double r = 1.0 / rb[1];
    ra[1] = ra[1] * r;
    rc[1] = rc[1] * r;
    rd[1] = rd[1] * r;
    A2[k] = ra[1];
    C2[k] = rc[1];
    D2[k] = rd[1];
    for (int i = 2; i < m; i++) {
      r = 1.0 / fma(-ra[i], rc[i - 1], rb[i]);
      rd[i] = fma(-ra[i], rd[i - 1], rd[i]) * r;
      ra[i] = -ra[i] * ra[i - 1] * r;
      rc[i] = rc[i] * r;
      A2[k] = ra[i];
      C2[k] = rc[i];
      D2[k] = rd[i];
    }
    for (int i = m - 3; i >= 1; i--) {
      rd[i] = fma(-rc[i], rd[i + 1], rd[i]);
      ra[i] = fma(-rc[i], ra[i + 1], ra[i]);
      rc[i] = -rc[i] * rc[i + 1];
    }
    if (m == 2) {
      A1[k] = ra[0]; B1[k] = rb[0]; C1[k] = rc[0]; D1[k] = rd[0];
    } else {
      A1[k] = ra[0];
      B1[k] = fma(-rc[0], ra[1], rb[0]);
      C1[k] = -rc[0] * rc[1];
      D1[k] = fma(-rc[0], rd[1], rd[0]);
    }
  }
  for (int k = 0; k < L; k++) {
    ea[k] = 0.0; eb[k] = 1.0; ec[k] = 0.0; ed[k] = 0.0;
    if (k >= q.p) continue;
    double a1n = 0.0, b1n = 1.0, c1n = 0.0, d1n = 0.0;
    if (k + 1 < L) {
      a1n = A1[k + 1]; b1n = B1[k + 1]; c1n = C1[k + 1]; d1n = D1[k + 1];
    }
    const double i1 = 1.0 / B1[k], in = 1.0 / b1n;
    ea[k] = -A2[k] * A1[k] * i1;
    eb[k] = 1.0 - A2[k] * C1[k] * i1 - C2[k] * a1n * in;
    ec[k] = -C2[k] * c1n * in;
    ed[k] = D2[k] - A2[k] * D1[k] * i1 - C2[k] * d1n * in;
  }
  for (int st = 1; st < L; st <<= 1) {
    double na[LMAX], nb[LMAX], nc[LMAX], nd[LMAX];
    for (int k = 0; k < L; k++) {
      double am = 0.0, bm = 1.0, cm = 0.0, dm = 0.0, ap = 0.0, bp = 1.0, cp = 0.0, dp = 0.0;
      if (k >= st) { am = ea[k - st]; bm = eb[k - st]; cm = ec[k - st]; dm = ed[k - st]; }
      if (k + st < L) { ap = ea[k + st]; bp = eb[k + st]; cp = ec[k + st]; dp = ed[k + st]; }
      const double k1 = ea[k] * (1.0 / bm), k2 = ec[k] * (1.0 / bp);
      na[k] = -am * k1;
      nc[k] = -cp * k2;
      nb[k] = fma(-ap, k2, fma(-cm, k1, eb[k]));
      nd[k] = fma(-dp, k2, fma(-dm, k1, ed[k]));
    }
    memcpy(ea, na, sizeof(double) * L);
    memcpy(eb, nb, sizeof(double) * L);
    memcpy(ec, nc, sizeof(double) * L);
    memcpy(ed, nd, sizeof(double) * L);
  }
  for (int k = 0; k < L; k++) xe[k] = ed[k] * (1.0 / eb[k]);
  U[0] = x0;
  U[n - 1] = xn;
  for (int k = 0; k < q.p; k++) {
    const int s = 1 + k * q.m, m = k == q.p - 1 && !full ? q.mlast : q.m;
    const int mreal = k == q.p - 1 ? q.mlast : q.m;
    const double xp = k == 0 ? 0.0 : xe[k - 1];
    const double xs = fma(-C1[k], xe[k], fma(-A1[k], xp, D1[k])) * (1.0 / B1[k]);
    U[s] = xs;
    for (int i = 1; i < mreal; i++)
      U[s + i] = i == m - 1 ? xe[k] : fma(-c[s + i], xe[k], fma(-a[s + i], xs, d[s + i]));
  }
  free(w);
  return 0;
}

static int g_line_solver = MTS_ORACLE_SOLVE_THOMAS;

void mts_oracle_set_line_solver(int kind) { g_line_solver = kind; }

int mts_oracle_line_solve(int n, const double *a, const double *b, const double *c,
                          const double *r, double *u) {
  if (n < 2) return -1;
  if (g_line_solver == MTS_ORACLE_SOLVE_PIVOT) return gtsv(n, a, b, c, r, u);
  if (g_line_solver == MTS_ORACLE_SOLVE_PARTITIONED) return part_solve(n, a, b, c, r, u);
  return mts_oracle_tdma(n, a, b, c, r, u);
}

/* Convert2Tri_irr (douglas_adi.cc:1554-1577) */
static void convert_irr(int ix, double row[2][5], double *a, double *b, double *c, double *r) {
  double rate = row[1][0] / row[0][0];
  for (int i = 0; i < 5; i++) row[1][i] -= rate * row[0][i];
  rate = row[0][3] / row[1][3];
  for (int i = 0; i < 5; i++) row[0][i] -= rate * row[1][i];
  for (int i = 0; i < 2; i++) {
    a[ix + i] = row[i][i];
    b[ix + i] = row[i][i + 1];
    c[ix + i] = row[i][i + 2];
    r[ix + i] = row[i][4];
  }
}

/* Convert2Tri_cor (douglas_adi.cc:1637-1675) */
static void convert_cor(int ix, double row[3][6], double *a, double *b, double *c, double *r) {
  double rate[3];
  rate[1] = row[1][0] / row[0][0];
  rate[2] = row[2][0] / row[0][0];
  for (int i = 0; i < 6; i++) {
    row[1][i] -= rate[1] * row[0][i];
    row[2][i] -= rate[2] * row[0][i];
  }
  rate[0] = row[0][4] / row[2][4];
  rate[1] = row[1][4] / row[2][4];
  for (int i = 0; i < 6; i++) {
    row[0][i] -= rate[0] * row[2][i];
    row[1][i] -= rate[1] * row[2][i];
  }
  rate[2] = row[2][1] / row[1][1];
  for (int i = 0; i < 6; i++) row[2][i] -= rate[2] * row[1][i];
  rate[0] = row[0][3] / row[1][3];
  for (int i = 0; i < 6; i++) row[0][i] -= rate[0] * row[1][i];
  for (int i = 0; i < 3; i++) {
    a[ix + i] = row[i][i];
    b[ix + i] = row[i][i + 1];
    c[ix + i] = row[i][i + 2];
    r[ix + i] = row[i][5];
  }
}

/* --------------------------------------------------------- field helpers */
/* Set_bc (lod.cc:216-236) / Douglas "Step V" (douglas_adi.cc:342-359) */
static void set_bc(const Prob *P, double *v, double e) {
  const int nx = P->n[0], ny = P->n[1], nz = P->n[2];
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++)
        if (ix == 0 || iy == 0 || iz == 0 || ix == nx - 1 || iy == ny - 1 || iz == nz - 1)
          v[IDX(P, ix, iy, iz)] = u_out(e, P->c[0][ix], P->c[1][iy], P->c[2][iz]);
}

/* Src_2nd source part (douglas_adi.cc:1137-1156, lod.cc:1509-1531) */
static void source_field(const Prob *P, double e, double *src) {
  for (int iz = 0; iz < P->n[2]; iz++)
    for (int iy = 0; iy < P->n[1]; iy++)
      for (int ix = 0; ix < P->n[0]; ix++) {
        double x = P->c[0][ix], y = P->c[1][iy], z = P->c[2][iz];
        src[IDX(P, ix, iy, iz)] = P->ind[IDX(P, ix, iy, iz)] > 0 ? f_out(P, e, x, y, z)
                                                                : f_in(P, e, x, y, z);
      }
}

/* Douglas Src_2nd jump part (douglas_adi.cc:1158-1428), directions x, y, z */
static void source_jumps(const Prob *P, double *src) {
  for (int d = 0; d < 3; d++) {
    int nl = P->nj[d] * P->nk[d];
    for (int lid = 0; lid < nl; lid++) {
      int j = lid % P->nj[d], k = lid / P->nj[d];
      int ip = P->lptr[d][lid], end = P->lptr[d][lid + 1];
      while (ip < end) {
        const Cr *D = &P->cr[d][ip];
        double vd[3];
        if (D->id > 0) {
          opw_row(P, d, j, k, D->ll, vd);
          src[NODE(P, d, j, k, D->ll)] += vd[2] * (D->ju * D->wr[4] + D->jb * D->wr[5]);
          opw_row(P, d, j, k, D->ll + 1, vd);
          src[NODE(P, d, j, k, D->ll + 1)] += vd[0] * (D->ju * D->wl[4] + D->jb * D->wl[5]);
          ip += 1;
        } else {
          const Cr *R = &P->cr[d][ip + 1];
          opw_row(P, d, j, k, D->ll, vd);
          src[NODE(P, d, j, k, D->ll)] += vd[2] * JL(D->wr, D, R);
          opw_row(P, d, j, k, D->ll + 1, vd);
          src[NODE(P, d, j, k, D->ll + 1)] += vd[0] * JL(D->wl, D, R);
          src[NODE(P, d, j, k, D->ll + 1)] += vd[2] * JL(R->wr, D, R);
          opw_row(P, d, j, k, D->ll + 2, vd);
          src[NODE(P, d, j, k, D->ll + 2)] += vd[0] * JL(R->wl, D, R);
          ip += 2;
        }
      }
    }
  }
}

/* Douglas "coef * L_d F" with fictitious points and no jump terms
 * (douglas_adi.cc:93-201 y, :205-312 z, :524-641, :817-934) */
static void douglas_T(const Prob *P, int d, const double *F, double *out, double coef) {
  const int nx = P->n[0], ny = P->n[1], nz = P->n[2];
  size_t st = d == 0 ? 1 : d == 1 ? (size_t)nx : (size_t)nx * ny;
  for (int iz = 1; iz < nz - 1; iz++)
    for (int iy = 1; iy < ny - 1; iy++)
      for (int ix = 1; ix < nx - 1; ix++) {
        size_t c = IDX(P, ix, iy, iz);
        double vd[3], sum = 0;
        opw(P, d, ix, iy, iz, vd);
        sum += F[c - st] * vd[0];
        sum += F[c] * vd[1];
        sum += F[c + st] * vd[2];
        out[c] = coef * sum;
      }
  int nl = P->nj[d] * P->nk[d];
  for (int lid = 0; lid < nl; lid++) {
    int j = lid % P->nj[d], k = lid / P->nj[d];
    int ip = P->lptr[d][lid], end = P->lptr[d][lid + 1];
#define FV(i) F[NODE(P, d, j, k, (i))]
    while (ip < end) {
      const Cr *D = &P->cr[d][ip];
      double vd[3], sum;
      int ll = D->ll;
      if (D->id > 0) {
        opw_row(P, d, j, k, ll, vd);
        sum = 0;
        for (int i = -1; i < 1; i++) sum += vd[i + 1] * FV(ll + i);
        for (int i = 0; i < 4; i++) sum += vd[2] * FV(ll - 1 + i) * D->wr[i];
        out[NODE(P, d, j, k, ll)] = sum * coef;
        opw_row(P, d, j, k, ll + 1, vd);
        sum = 0;
        for (int i = 0; i < 2; i++) sum += vd[i + 1] * FV(ll + 1 + i);
        for (int i = 0; i < 4; i++) sum += vd[0] * FV(ll - 1 + i) * D->wl[i];
        out[NODE(P, d, j, k, ll + 1)] = sum * coef;
        ip += 1;
      } else {
        const Cr *R = &P->cr[d][ip + 1];
        opw_row(P, d, j, k, ll, vd);
        sum = 0;
        for (int i = -1; i < 1; i++) sum += vd[i + 1] * FV(ll + i);
        for (int i = 0; i < 5; i++) sum += vd[2] * FV(ll - 1 + i) * D->wr[i];
        out[NODE(P, d, j, k, ll)] = sum * coef;
        opw_row(P, d, j, k, ll + 1, vd);
        sum = 0;
        for (int i = 0; i < 5; i++) sum += vd[0] * FV(ll - 1 + i) * D->wl[i];
        sum += vd[1] * FV(ll + 1);
        for (int i = 0; i < 5; i++) sum += vd[2] * FV(ll - 1 + i) * R->wr[i];
        out[NODE(P, d, j, k, ll + 1)] = sum * coef;
        opw_row(P, d, j, k, ll + 2, vd);
        sum = 0;
        for (int i = 0; i < 2; i++) sum += vd[i + 1] * FV(ll + 2 + i);
        for (int i = 0; i < 5; i++) sum += vd[0] * FV(ll - 1 + i) * R->wl[i];
        out[NODE(P, d, j, k, ll + 2)] = sum * coef;
        ip += 2;
      }
    }
#undef FV
  }
}

/* (1 + coef dt beta D_dd) F with fictitious points and jump terms
 * (LOD::D_xx_r_cn lod.cc:309-485 and its y/z twins; TS::D_xx_r
 * trapezoidal.cc:319-878): interior nodes only. */
static void cn_rhs(const Prob *P, int d, const double *F, double *out, double coef, double dt) {
  const int nx = P->n[0], ny = P->n[1], nz = P->n[2];
  size_t st = d == 0 ? 1 : d == 1 ? (size_t)nx : (size_t)nx * ny;
  for (int iz = 1; iz < nz - 1; iz++)
    for (int iy = 1; iy < ny - 1; iy++)
      for (int ix = 1; ix < nx - 1; ix++) {
        size_t c = IDX(P, ix, iy, iz);
        double vd[3], sum = 0;
        opw(P, d, ix, iy, iz, vd);
        sum += F[c - st] * vd[0];
        sum += F[c] * vd[1];
        sum += F[c + st] * vd[2];
        out[c] = coef * dt * sum;
      }
  int nl = P->nj[d] * P->nk[d];
  for (int lid = 0; lid < nl; lid++) {
    int j = lid % P->nj[d], k = lid / P->nj[d];
    int ip = P->lptr[d][lid], end = P->lptr[d][lid + 1];
#define FV(i) F[NODE(P, d, j, k, (i))]
    while (ip < end) {
      const Cr *D = &P->cr[d][ip];
      double vd[3], sum;
      int ll = D->ll;
      if (D->id > 0) {
        opw_row(P, d, j, k, ll, vd);
        sum = 0;
        for (int i = -1; i < 1; i++) sum += vd[i + 1] * FV(ll + i);
        for (int i = 0; i < 4; i++) sum += vd[2] * FV(ll - 1 + i) * D->wr[i];
        sum += vd[2] * (D->ju * D->wr[4] + D->jb * D->wr[5]);
        out[NODE(P, d, j, k, ll)] = coef * dt * sum;
        opw_row(P, d, j, k, ll + 1, vd);
        sum = 0;
        for (int i = 0; i < 2; i++) sum += vd[i + 1] * FV(ll + 1 + i);
        for (int i = 0; i < 4; i++) sum += vd[0] * FV(ll - 1 + i) * D->wl[i];
        sum += vd[0] * (D->ju * D->wl[4] + D->jb * D->wl[5]);
        out[NODE(P, d, j, k, ll + 1)] = coef * dt * sum;
        ip += 1;
      } else {
        const Cr *R = &P->cr[d][ip + 1];
        opw_row(P, d, j, k, ll, vd);
        sum = 0;
        for (int i = -1; i < 1; i++) sum += vd[i + 1] * FV(ll + i);
        for (int i = 0; i < 5; i++) sum += vd[2] * FV(ll - 1 + i) * D->wr[i];
        sum += vd[2] * JL(D->wr, D, R);
        out[NODE(P, d, j, k, ll)] = coef * dt * sum;
        opw_row(P, d, j, k, ll + 1, vd);
        sum = 0;
        for (int i = 0; i < 5; i++) sum += vd[0] * FV(ll - 1 + i) * D->wl[i];
        sum += vd[0] * JL(D->wl, D, R);
        sum += vd[1] * FV(ll + 1);
        for (int i = 0; i < 5; i++) sum += vd[2] * FV(ll - 1 + i) * R->wr[i];
        sum += vd[2] * JL(R->wr, D, R);
        out[NODE(P, d, j, k, ll + 1)] = coef * dt * sum;
        opw_row(P, d, j, k, ll + 2, vd);
        sum = 0;
        for (int i = 0; i < 2; i++) sum += vd[i + 1] * FV(ll + 2 + i);
        for (int i = 0; i < 5; i++) sum += vd[0] * FV(ll - 1 + i) * R->wl[i];
        sum += vd[0] * JL(R->wl, D, R);
        out[NODE(P, d, j, k, ll + 2)] = coef * dt * sum;
        ip += 2;
      }
    }
#undef FV
  }
  for (int iz = 1; iz < nz - 1; iz++)
    for (int iy = 1; iy < ny - 1; iy++)
      for (int ix = 1; ix < nx - 1; ix++) out[IDX(P, ix, iy, iz)] += F[IDX(P, ix, iy, iz)];
}

/* Implicit sweep along d on every interior line of R (in place):
 * Douglas LHS loops (douglas_adi.cc:377-514, 670-807, 963-1100) when jumps == 0,
 * LOD::D_xx_l / TS::D_xx_l (lod.cc:887-1078) when jumps == 1. */
static int sweep(Prob *P, int d, double *R, double coef, double dt, int jumps) {
  const int n = P->n[d];
  double *a = P->la, *b = P->lb, *c = P->lc, *r = P->lr, *ut = P->lu;
  const double cdt = coef * dt;
  for (int k = 1; k < P->nk[d] - 1; k++)
    for (int j = 1; j < P->nj[d] - 1; j++) {
      int lid = k * P->nj[d] + j;
      double vd[3];
      a[0] = 0.0;
      b[0] = 1.0;
      c[0] = 0.0;
      for (int i = 1; i < n - 1; i++) {
        opw_row(P, d, j, k, i, vd);
        a[i] = -coef * dt * vd[0];
        b[i] = 1 - coef * dt * vd[1];
        c[i] = -coef * dt * vd[2];
      }
      a[n - 1] = 0.0;
      b[n - 1] = 1.0;
      c[n - 1] = 0.0;
      for (int i = 0; i < n; i++) r[i] = R[NODE(P, d, j, k, i)];
      int ip = P->lptr[d][lid], end = P->lptr[d][lid + 1];
      while (ip < end) {
        const Cr *D = &P->cr[d][ip];
        int ll = D->ll;
        if (D->id > 0) {
          double row[2][5];
          memset(row, 0, sizeof(row));
          opw_row(P, d, j, k, ll, vd);
          row[0][0] = -coef * dt * vd[0];
          row[0][1] = 1 - coef * dt * vd[1];
          for (int i = 0; i < 4; i++) row[0][i] += -coef * dt * vd[2] * D->wr[i];
          row[0][4] = R[NODE(P, d, j, k, ll)];
          if (jumps) row[0][4] += cdt * vd[2] * (D->ju * D->wr[4] + D->jb * D->wr[5]);
          opw_row(P, d, j, k, ll + 1, vd);
          row[1][2] = 1 - coef * dt * vd[1];
          row[1][3] = -coef * dt * vd[2];
          for (int i = 0; i < 4; i++) row[1][i] += -coef * dt * vd[0] * D->wl[i];
          row[1][4] = R[NODE(P, d, j, k, ll + 1)];
          if (jumps) row[1][4] += cdt * vd[0] * (D->ju * D->wl[4] + D->jb * D->wl[5]);
          convert_irr(ll, row, a, b, c, r);
          ip += 1;
        } else {
          const Cr *Rr = &P->cr[d][ip + 1];
          double row[3][6];
          memset(row, 0, sizeof(row));
          opw_row(P, d, j, k, ll, vd);
          row[0][0] = -coef * dt * vd[0];
          row[0][1] = 1 - coef * dt * vd[1];
          for (int i = 0; i < 5; i++) row[0][i] -= coef * dt * vd[2] * D->wr[i];
          row[0][5] = R[NODE(P, d, j, k, ll)];
          if (jumps) row[0][5] += cdt * vd[2] * JL(D->wr, D, Rr);
          opw_row(P, d, j, k, ll + 1, vd);
          row[1][2] = 1 - coef * dt * vd[1];
          for (int i = 0; i < 5; i++) row[1][i] -= coef * dt * vd[0] * D->wl[i];
          row[1][5] = R[NODE(P, d, j, k, ll + 1)];
          if (jumps) row[1][5] += cdt * vd[0] * JL(D->wl, D, Rr);
          for (int i = 0; i < 5; i++) row[1][i] -= coef * dt * vd[2] * Rr->wr[i];
          if (jumps) row[1][5] += cdt * vd[2] * JL(Rr->wr, D, Rr);
          opw_row(P, d, j, k, ll + 2, vd);
          row[2][3] = 1 - coef * dt * vd[1];
          row[2][4] = -coef * dt * vd[2];
          for (int i = 0; i < 5; i++) row[2][i] -= coef * dt * vd[0] * Rr->wl[i];
          row[2][5] = R[NODE(P, d, j, k, ll + 2)];
          if (jumps) row[2][5] += cdt * vd[0] * JL(Rr->wl, D, Rr);
          convert_cor(ll, row, a, b, c, r);
          ip += 2;
        }
      }
      if (mts_oracle_line_solve(n, a, b, c, r, ut) != 0) return -1;
      for (int i = 0; i < n; i++) R[NODE(P, d, j, k, i)] = ut[i];
    }
  return 0;
}

/* --------------------------------------------------------------- problem */
static void prob_free(Prob *P) {
  for (int a = 0; a < 3; a++) {
    free(P->c[a]);
    free(P->cr[a]);
    free(P->lptr[a]);
  }
  free(P->la);
  free(P->lb);
  free(P->lc);
  free(P->lr);
  free(P->lu);
  free(P->gam);
}

static int prob_init(Prob *P, const mts_oracle_problem *p) {
  memset(P, 0, sizeof(*P));
  int nmax = 0;
  for (int a = 0; a < 3; a++) {
    P->n[a] = p->n[a];
    if (p->n[a] < 3) return -1;
    P->h[a] = (p->hi[a] - p->lo[a]) / (p->n[a] - 1.0);
    P->c[a] = (double *)malloc(sizeof(double) * (size_t)p->n[a]);
    if (!P->c[a]) return -1;
    for (int i = 0; i < p->n[a]; i++) P->c[a][i] = p->lo[a] + i * P->h[a];
    if (p->n[a] > nmax) nmax = p->n[a];
  }
  P->N = (size_t)P->n[0] * P->n[1] * P->n[2];
  P->ind = p->ind;
  P->bk = p->beta_kind;
  P->b0 = p->bp[0];
  P->b1 = p->bp[1];
  for (int d = 0; d < 3; d++) {
    int ja, ka;
    axes(d, &ja, &ka);
    P->nj[d] = P->n[ja];
    P->nk[d] = P->n[ka];
    int nl = P->nj[d] * P->nk[d];
    P->ncr[d] = p->ncross[d];
    P->cr[d] = (Cr *)calloc((size_t)(p->ncross[d] > 0 ? p->ncross[d] : 1), sizeof(Cr));
    P->lptr[d] = (int *)calloc((size_t)nl + 1, sizeof(int));
    if (!P->cr[d] || !P->lptr[d]) return -1;
    int prev = -1;
    for (int i = 0; i < p->ncross[d]; i++) {
      Cr *c = &P->cr[d][i];
      const int *ii = p->cr_int[d] + 4 * (size_t)i;
      const double *dd = p->cr_dbl[d] + 24 * (size_t)i;
      c->j = ii[0];
      c->k = ii[1];
      c->id = ii[2];
      c->ll = ii[3];
      memcpy(c->coord, dd, 3 * sizeof(double));
      memcpy(c->normal, dd + 3, 3 * sizeof(double));
      memcpy(c->wl, dd + 6, 9 * sizeof(double));
      memcpy(c->wr, dd + 15, 9 * sizeof(double));
      int lid = c->k * P->nj[d] + c->j;
      if (lid < prev || lid < 0 || lid >= nl) return -1;
      prev = lid;
      P->lptr[d][lid + 1]++;
    }
    for (int l = 0; l < nl; l++) P->lptr[d][l + 1] += P->lptr[d][l];
  }
  P->la = (double *)malloc(sizeof(double) * nmax);
  P->lb = (double *)malloc(sizeof(double) * nmax);
  P->lc = (double *)malloc(sizeof(double) * nmax);
  P->lr = (double *)malloc(sizeof(double) * nmax);
  P->lu = (double *)malloc(sizeof(double) * nmax);
  if (!P->la || !P->lb || !P->lc || !P->lr || !P->lu) return -1;
  return 0;
}

int mts_oracle_init(const mts_oracle_problem *p, double t0, double *u) {
  Prob P;
  if (prob_init(&P, p)) {
    prob_free(&P);
    return -1;
  }
  double e = exp(-t0);
  for (int iz = 0; iz < P.n[2]; iz++)
    for (int iy = 0; iy < P.n[1]; iy++)
      for (int ix = 0; ix < P.n[0]; ix++) {
        double x = P.c[0][ix], y = P.c[1][iy], z = P.c[2][iz];
        size_t c = IDX(&P, ix, iy, iz);
        u[c] = P.ind[c] > 0 ? u_out(e, x, y, z) : u_in(e, x, y, z);
      }
  prob_free(&P);
  return 0;
}

/* Douglas_ADI::Solve_2nd (douglas_adi.cc:66-1110).  stages (may be null):
 * receives the stage-1 right-hand side with its boundary values (the input
 * of the x sweep), u* and u** (the outputs of the x and y sweeps). */
static int step_adi_stages(Prob *P, double dt, double t1, double *u, double *w[5],
                           double *stages[3]) {
  double *uhs = w[0], *uhs2 = w[1], *uhss = w[2], *uh1 = w[3], *src = w[4];
  const double e = exp(-t1);
  const int nx = P->n[0], ny = P->n[1], nz = P->n[2];
  refresh_jumps(P, e);
  source_field(P, e, src);
  source_jumps(P, src);
  douglas_T(P, 1, u, uhs, dt);
  douglas_T(P, 2, u, uhs2, dt);
  for (int iz = 1; iz < nz - 1; iz++)
    for (int iy = 1; iy < ny - 1; iy++)
      for (int ix = 1; ix < nx - 1; ix++) {
        size_t c = IDX(P, ix, iy, iz);
        uhs[c] += uhs2[c];
        uhs[c] += u[c];
        uhs[c] += dt * src[c];
      }
  set_bc(P, uhs, e);
  if (stages && stages[0]) memcpy(stages[0], uhs, P->N * sizeof(double));
  if (sweep(P, 0, uhs, 1.0, dt, 0)) return -1;
  if (stages && stages[1]) memcpy(stages[1], uhs, P->N * sizeof(double));
  douglas_T(P, 1, u, uhss, -dt);
  for (int iz = 1; iz < nz - 1; iz++)
    for (int iy = 1; iy < ny - 1; iy++)
      for (int ix = 1; ix < nx - 1; ix++) uhss[IDX(P, ix, iy, iz)] += uhs[IDX(P, ix, iy, iz)];
  set_bc(P, uhss, e);
  if (sweep(P, 1, uhss, 1.0, dt, 0)) return -1;
  if (stages && stages[2]) memcpy(stages[2], uhss, P->N * sizeof(double));
  douglas_T(P, 2, u, uh1, -dt);
  for (int iz = 1; iz < nz - 1; iz++)
    for (int iy = 1; iy < ny - 1; iy++)
      for (int ix = 1; ix < nx - 1; ix++) uh1[IDX(P, ix, iy, iz)] += uhss[IDX(P, ix, iy, iz)];
  set_bc(P, uh1, e);
  if (sweep(P, 2, uh1, 1.0, dt, 0)) return -1;
  memcpy(u, uh1, P->N * sizeof(double));
  return 0;
}

static int step_adi(Prob *P, double dt, double t1, double *u, double *w[5]) {
  return step_adi_stages(P, dt, t1, u, w, NULL);
}

/* LOD::Solve_2nd (lod.cc:44-205); fresh zeroed stage fields each step (lod.cc:53-71) */
static int step_lod(Prob *P, double dt, double t1, double *u, double *w[5], int cn) {
  double *v1 = w[0], *v2 = w[1], *v3 = w[2], *src = w[4];
  const double e = exp(-t1);
  const double coef = cn ? 0.5 : 1.0;
  refresh_jumps(P, e);
  source_field(P, e, src);
  double *in = u;
  double *vs[3] = {v1, v2, v3};
  for (int d = 0; d < 3; d++) {
    double *v = vs[d];
    memset(v, 0, P->N * sizeof(double));
    if (cn) cn_rhs(P, d, in, v, coef, dt);
    else memcpy(v, in, P->N * sizeof(double));
    set_bc(P, v, e);
    if (sweep(P, d, v, coef, dt, 1)) return -1;
    in = v;
  }
  for (size_t i = 0; i < P->N; i++) u[i] = v3[i] + dt * src[i];
  return 0;
}

/* TS::Solve_2nd (trapezoidal.cc:46-272) */
static int step_ts(Prob *P, double dt, double t1, double *u, double *w[5]) {
  double *va = w[0], *vb = w[1], *src = w[4];
  const double en = exp(-(t1 - dt)), eh = exp(-(t1 - 0.5 * dt)), e1 = exp(-t1);
  refresh_jumps(P, e1);
  /* v1 = (1 + dt/2 D_xx) u, v2 = (1 + dt/2 D_yy) v1, v3 = (1 + dt/2 D_zz) v2 */
  memset(va, 0, P->N * sizeof(double));
  cn_rhs(P, 0, u, va, 0.5, dt);
  set_bc(P, va, eh);
  memset(vb, 0, P->N * sizeof(double));
  cn_rhs(P, 1, va, vb, 0.5, dt);
  set_bc(P, vb, eh);
  memset(va, 0, P->N * sizeof(double));
  cn_rhs(P, 2, vb, va, 0.5, dt);
  set_bc(P, va, eh);
  /* v4 = v3 + dt/2 f(t^n), v5 = v4 + dt/2 f(t^{n+1}) */
  source_field(P, en, src);
  for (size_t i = 0; i < P->N; i++) va[i] = va[i] + 0.5 * dt * src[i];
  source_field(P, e1, src);
  for (size_t i = 0; i < P->N; i++) va[i] = va[i] + 0.5 * dt * src[i];
  /* implicit z, y, x */
  set_bc(P, va, e1);
  if (sweep(P, 2, va, 0.5, dt, 1)) return -1;
  set_bc(P, va, e1);
  if (sweep(P, 1, va, 0.5, dt, 1)) return -1;
  set_bc(P, va, e1);
  if (sweep(P, 0, va, 0.5, dt, 1)) return -1;
  memcpy(u, va, P->N * sizeof(double));
  return 0;
}

int mts_oracle_run(const mts_oracle_problem *p, int method, double dt, double t0, int nsteps,
                   double *u, double *t_end) {
  Prob P;
  if (prob_init(&P, p)) {
    prob_free(&P);
    return -1;
  }
  double *w[5];
  int rc = 0;
  for (int i = 0; i < 5; i++) {
    w[i] = (double *)calloc(P.N, sizeof(double));
    if (!w[i]) rc = -1;
  }
  double t = t0;
  for (int s = 0; s < nsteps && rc == 0; s++) {
    t += dt;
    if (method == MTS_ORACLE_ADI) rc = step_adi(&P, dt, t, u, w);
    else if (method == MTS_ORACLE_LOD_IE) rc = step_lod(&P, dt, t, u, w, 0);
    else if (method == MTS_ORACLE_LOD_CN) rc = step_lod(&P, dt, t, u, w, 1);
    else if (method == MTS_ORACLE_TS) rc = step_ts(&P, dt, t, u, w);
    else rc = -1;
  }
  if (t_end) *t_end = t;
  for (int i = 0; i < 5; i++) free(w[i]);
  prob_free(&P);
  return rc;
}

int mts_oracle_adi_stages(const mts_oracle_problem *p, double dt, double t0, double *u,
                          double *rhs1, double *ustar, double *ustar2) {
  Prob P;
  if (prob_init(&P, p)) {
    prob_free(&P);
    return -1;
  }
  double *w[5];
  int rc = 0;
  for (int i = 0; i < 5; i++) {
    w[i] = (double *)calloc(P.N, sizeof(double));
    if (!w[i]) rc = -1;
  }
  double *stages[3] = {rhs1, ustar, ustar2};
  if (rc == 0) rc = step_adi_stages(&P, dt, t0 + dt, u, w, stages);
  for (int i = 0; i < 5; i++) free(w[i]);
  prob_free(&P);
  return rc;
}

void mts_oracle_norms(const mts_oracle_problem *p, const double *u, double t, double *linf,
                      double *l2, int argmax[3]) {
  Prob P;
  double lmax = 0.0, s2 = 0.0;
  int xm = 0, ym = 0, zm = 0;
  if (prob_init(&P, p)) {
    prob_free(&P);
    *linf = -1.0;
    *l2 = -1.0;
    return;
  }
  double e = exp(-t);
  for (int ix = 0; ix < P.n[0]; ix++)
    for (int iy = 0; iy < P.n[1]; iy++)
      for (int iz = 0; iz < P.n[2]; iz++) {
        size_t c = IDX(&P, ix, iy, iz);
        double x = P.c[0][ix], y = P.c[1][iy], z = P.c[2][iz];
        double ex = P.ind[c] > 0 ? u_out(e, x, y, z) : u_in(e, x, y, z);
        double er = fabs(u[c] - ex);
        if (er > lmax) {
          lmax = er;
          xm = ix;
          ym = iy;
          zm = iz;
        }
        s2 += er * er;
      }
  *linf = lmax;
  *l2 = sqrt(s2 / (1.0 * P.n[0] * P.n[1] * P.n[2]));
  argmax[0] = xm;
  argmax[1] = ym;
  argmax[2] = zm;
  prob_free(&P);
}
