// Matrix part of the partitioned line solve for regular lines (see
// regular_solve.hpp): the line kernels' group_solve run on the coefficients
// alone, over every lane of a group, in FP64.
#include "regular_solve.hpp"

#include <cstring>

#include "device_core.hpp"

namespace mts {

bool regular_factors(const LinePart &q, int mmax, bool full, double a, double b, double c,
                     std::vector<double> &out) {
  const int L = q.L;
  if (mmax < 2 || L < 1 || L > 64 || (L & (L - 1))) return false;
  int levels = 0;
  while ((1 << levels) < L) levels++;
  if (levels > kRegMaxLevels) return false;
  const size_t base = out.size();
  out.resize(base + (size_t)reg_side_len(L, mmax), 0.0);
  double *tab = out.data() + base;

  struct Lane {
    std::vector<double> A, B, C, r, aorig, cf;
    double A1 = 0, B1 = 1, C1 = 0, A2 = 0, C2 = 0, cD = 0;
    int m = 0, meff = 0;
  };
  std::vector<Lane> ln((size_t)L);
  for (int k = 0; k < L; k++) {
    Lane &l = ln[(size_t)k];
    l.m = k < q.p ? (k == q.p - 1 ? q.mlast : q.m) : 0;
    l.meff = full ? mmax : l.m;
    l.A.assign((size_t)mmax, 0.0);
    l.B.assign((size_t)mmax, 1.0);
    l.C.assign((size_t)mmax, 0.0);
    l.r.assign((size_t)mmax, 0.0);
    l.aorig.assign((size_t)mmax, 0.0);
    l.cf.assign((size_t)mmax, 0.0);
    for (int i = 0; i < l.m; i++) {
      l.A[(size_t)i] = (k == 0 && i == 0) ? 0.0 : a;  // first row: coupling moved to the RHS
      l.B[(size_t)i] = b;
      l.C[(size_t)i] = (k == q.p - 1 && i == l.m - 1) ? 0.0 : c;  // last row likewise
    }
    l.aorig = l.A;
    const int m = l.meff;
    if (m < 2) continue;
    std::vector<double> &A = l.A, &B = l.B, &C = l.C;
    // downward elimination keeping the first unknown as a parameter
    {
      if (B[1] == 0.0) return false;
      const double r = 1.0 / B[1];
      l.r[1] = r;
      A[1] = A[1] * r;
      C[1] = C[1] * r;
      l.A2 = A[1];
      l.C2 = C[1];
    }
    for (int i = 2; i < mmax; i++) {
      if (!(full || i < m)) continue;
      const double piv = B[(size_t)i] - l.aorig[(size_t)i] * C[(size_t)i - 1];
      if (piv == 0.0) return false;
      const double r = 1.0 / piv;
      l.r[(size_t)i] = r;
      A[(size_t)i] = -l.aorig[(size_t)i] * A[(size_t)i - 1] * r;
      C[(size_t)i] = C[(size_t)i] * r;
      l.A2 = A[(size_t)i];
      l.C2 = C[(size_t)i];
    }
    l.cf = C;  // normalized super-diagonal after the downward pass
    // upward elimination
    for (int i = mmax - 3; i >= 1; i--) {
      if (!(full || i <= m - 3)) continue;
      A[(size_t)i] = A[(size_t)i] - C[(size_t)i] * A[(size_t)i + 1];
      C[(size_t)i] = -C[(size_t)i] * C[(size_t)i + 1];
    }
    // the segment's first row
    if (m == 2) {
      l.A1 = l.aorig[0];
      l.B1 = B[0];
      l.C1 = C[0];
      l.cD = 0.0;
    } else {
      l.A1 = l.aorig[0];
      l.B1 = B[0] - C[0] * A[1];
      l.C1 = -C[0] * C[1];
      l.cD = C[0];
    }
    if (l.B1 == 0.0) return false;
  }
  // fold and cyclic reduction over the group
  std::vector<double> ea((size_t)L), eb((size_t)L), ec((size_t)L);
  std::vector<std::vector<double>> k1((size_t)levels, std::vector<double>((size_t)L)),
      k2((size_t)levels, std::vector<double>((size_t)L));
  std::vector<double> p1((size_t)L, 0.0), p2((size_t)L, 0.0), i1((size_t)L);
  for (int k = 0; k < L; k++) {
    const Lane &l = ln[(size_t)k];
    const bool nxt = k + 1 < L;
    const double A1n = nxt ? ln[(size_t)k + 1].A1 : 0.0;
    const double B1n = nxt ? ln[(size_t)k + 1].B1 : 1.0;
    const double C1n = nxt ? ln[(size_t)k + 1].C1 : 0.0;
    i1[(size_t)k] = 1.0 / l.B1;
    const double in = 1.0 / B1n;
    ea[(size_t)k] = 0.0;
    eb[(size_t)k] = 1.0;
    ec[(size_t)k] = 0.0;
    if (l.meff >= 2) {
      ea[(size_t)k] = -l.A2 * l.A1 * i1[(size_t)k];
      eb[(size_t)k] = 1.0 - l.A2 * l.C1 * i1[(size_t)k] - l.C2 * A1n * in;
      ec[(size_t)k] = -l.C2 * C1n * in;
      p1[(size_t)k] = l.A2 * i1[(size_t)k];
      p2[(size_t)k] = l.C2 * in;
    }
  }
  for (int lv = 0; lv < levels; lv++) {
    const int s = 1 << lv;
    std::vector<double> na((size_t)L), nb((size_t)L), nc((size_t)L);
    for (int k = 0; k < L; k++) {
      const bool lo = k >= s, hi = k + s < L;
      const double am = lo ? ea[(size_t)(k - s)] : 0.0, bm = lo ? eb[(size_t)(k - s)] : 1.0;
      const double cm = lo ? ec[(size_t)(k - s)] : 0.0;
      const double ap = hi ? ea[(size_t)(k + s)] : 0.0, bp = hi ? eb[(size_t)(k + s)] : 1.0;
      const double cp = hi ? ec[(size_t)(k + s)] : 0.0;
      if (bm == 0.0 || bp == 0.0) return false;
      const double f1 = ea[(size_t)k] / bm, f2 = ec[(size_t)k] / bp;
      k1[(size_t)lv][(size_t)k] = f1;
      k2[(size_t)lv][(size_t)k] = f2;
      na[(size_t)k] = -am * f1;
      nc[(size_t)k] = -cp * f2;
      nb[(size_t)k] = eb[(size_t)k] - cm * f1 - ap * f2;
    }
    ea = na;
    eb = nb;
    ec = nc;
  }
  // the row factors of the common segment (lane 0; lanes 1 .. p-2 have the
  // same rows) and of the last segment
  for (int v = 0; v < 2; v++) {
    const int k = v == 0 ? 0 : q.p - 1;
    const Lane &l = ln[(size_t)k];
    double *rows = tab + reg_rows_off(v, mmax);
    for (int i = 0; i < mmax; i++) {
      double *w = rows + (size_t)i * kRegRow;
      w[kRegR] = l.r[(size_t)i];
      w[kRegA] = l.aorig[(size_t)i];
      w[kRegCf] = l.cf[(size_t)i];
      w[kRegAf] = l.A[(size_t)i];
      w[kRegCfin] = l.C[(size_t)i];
    }
  }
  for (int k = 1; k < q.p - 1; k++) {  // check: the middle segments share lane 0's rows
    const Lane &l = ln[(size_t)k], &l0 = ln[0];
    for (int i = 1; i < mmax; i++)
      if (l.r[(size_t)i] != l0.r[(size_t)i] || l.cf[(size_t)i] != l0.cf[(size_t)i] ||
          l.A[(size_t)i] != l0.A[(size_t)i] || l.C[(size_t)i] != l0.C[(size_t)i] ||
          l.aorig[(size_t)i] != l0.aorig[(size_t)i])
        return false;
  }
  for (int k = 0; k < L; k++) {
    const Lane &l = ln[(size_t)k];
    double *sc = tab + reg_scal_off(k, mmax);
    sc[kRegCD] = l.cD;
    sc[kRegI1] = i1[(size_t)k];
    sc[kRegA1] = l.A1;
    sc[kRegC1] = l.C1;
    sc[kRegP1] = p1[(size_t)k];
    sc[kRegP2] = p2[(size_t)k];
    for (int lv = 0; lv < levels; lv++) {
      sc[kRegK + 2 * lv] = k1[(size_t)lv][(size_t)k];
      sc[kRegK + 2 * lv + 1] = k2[(size_t)lv][(size_t)k];
    }
    sc[kRegQ] = 1.0 / eb[(size_t)k];
  }
  return true;
}

void tile_lists(const int *blk_ptr, const uint8_t *cls0, bool beta_const, int nj, int nk, int L,
                std::vector<int> &regular, std::vector<int> &general) {
  const int T = 256 / L;
  const int gx = (nj + T - 1) / T;
  regular.clear();
  general.clear();
  for (int k = 1; k <= nk - 2; k++)
    for (int bx = 0; bx < gx; bx++) {
      bool reg = true;
      for (int j = bx * T; j < bx * T + T && j < nj; j++) {
        if (j < 1 || j > nj - 2) continue;
        const int lid = k * nj + j;
        if (blk_ptr[lid] != blk_ptr[lid + 1] || !(cls0[lid] || beta_const)) reg = false;
      }
      (reg ? regular : general).push_back(k * 65536 + bx);
    }
}

void recut_tiles(const std::vector<int> &from, int t_from, int t_to, int nj, std::vector<int> &out) {
  out.clear();
  const int r = t_from / t_to;
  for (int e : from) {
    const int bx = tile_bx(e), k = tile_k(e);
    for (int i = 0; i < r; i++) {
      const int b2 = bx * r + i;
      if (b2 * t_to < nj) out.push_back(k * 65536 + b2);
    }
  }
}

void xcd_group_tiles(std::vector<int> &list, int group) {
  if (group <= 1) return;
  // line groups (same k, same bx / group) as they come in the list
  std::vector<std::vector<int>> full, part;
  for (size_t i = 0; i < list.size();) {
    const int key_k = tile_k(list[i]), key_b = tile_bx(list[i]) / group;
    std::vector<int> u;
    while (i < list.size() && tile_k(list[i]) == key_k && tile_bx(list[i]) / group == key_b)
      u.push_back(list[i++]);
    (u.size() == (size_t)group ? full : part).push_back(std::move(u));
  }
  std::vector<int> out;
  out.reserve(list.size());
  for (size_t u0 = 0; u0 < full.size(); u0 += 8) {
    const size_t nu = full.size() - u0 < 8 ? full.size() - u0 : 8;
    for (int m = 0; m < group; m++)
      for (size_t u = 0; u < nu; u++) out.push_back(full[u0 + u][(size_t)m]);
  }
  for (const auto &u : part) out.insert(out.end(), u.begin(), u.end());
  list.swap(out);
}

bool stage_tiles(const int *blk_ptr, const uint8_t *cls0, bool beta_const, int dir, int n, int nj,
                 int nk, int min_lanes, int gen_lanes, int elem_bytes, std::vector<int> &regular,
                 std::vector<int> &general) {
  LinePart q, qg;
  if (!make_line_part(n, q, min_lanes)) return false;
  if (!make_line_part(n, qg, gen_lanes > min_lanes ? gen_lanes : min_lanes)) return false;
  std::vector<int> gen;
  tile_lists(blk_ptr, cls0, beta_const, nj, nk, q.L, regular, gen);
  recut_tiles(gen, tile_lines(q.L), tile_lines(qg.L), nj, general);
  xcd_group_tiles(regular, line_group(dir, tile_lines(q.L), elem_bytes));
  xcd_group_tiles(general, line_group(dir, tile_lines(qg.L), elem_bytes));
  return true;
}

}  // namespace mts
