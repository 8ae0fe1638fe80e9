// Device table construction: blocks with precomputed row conversions.
#include "host_table.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>

#include "host_parallel.hpp"

namespace mts {

namespace {

inline void axes_of(int d, int &ja, int &ka) {
  if (d == 0) { ja = 1; ka = 2; }
  else if (d == 1) { ja = 0; ka = 2; }
  else { ja = 0; ka = 1; }
}

// node index triple of row i on line (d, j, k)
inline void node_of(int d, int j, int k, int i, int q[3]) {
  int ja, ka;
  axes_of(d, ja, ka);
  q[ja] = j;
  q[ka] = k;
  q[d] = i;
}

}  // namespace

void operator_weights(const SetupResult &s, const BetaModel &beta, int d, int ix, int iy,
                      int iz, double vd[3]) {
  const Grid &g = s.grid;
  double x = g.c[0][ix], y = g.c[1][iy], z = g.c[2][iz];
  double h = g.h[d];
  double p1[3] = {x, y, z}, p2[3] = {x, y, z};
  p1[d] = (d == 0 ? x : d == 1 ? y : z) - h / 2;
  p2[d] = (d == 0 ? x : d == 1 ? y : z) + h / 2;
  double b1, b2;
  if (s.ind[g.idx(ix, iy, iz)] > 0) {
    b1 = beta.outside(p1[0], p1[1], p1[2]);
    b2 = beta.outside(p2[0], p2[1], p2[2]);
  } else {
    b1 = beta.inside(p1[0], p1[1], p1[2]);
    b2 = beta.inside(p2[0], p2[1], p2[2]);
  }
  vd[0] = b1 / (h * h);
  vd[1] = -(b1 + b2) / (h * h);
  vd[2] = b2 / (h * h);
}

namespace {
// row (a, b, c) of a tridiagonal system is diagonally dominant
bool dominant(double a, double b, double c) { return !(std::fabs(b) < std::fabs(a) + std::fabs(c)); }

// Jump record of one crossing (kCrLen doubles).
void jump_record(const Crossing &c, double *r) {
  r[0] = std::sin(c.coord[0]);
  r[1] = std::sin(c.coord[1]);
  r[2] = std::sin(c.coord[2]);
  r[3] = std::cos(c.coord[0]);
  r[4] = std::cos(c.coord[1]);
  r[5] = std::cos(c.coord[2]);
  r[6] = c.beta_in;
  r[7] = c.beta_out;
  r[8] = c.p[0][0];
  r[9] = c.p[0][1];
  r[10] = c.p[0][2];
  r[11] = c.pinv_d0;
}

// Record of the block h on line (d, j, k) into rec (zeroed, kBlkIrrLen or
// kBlkCorLen doubles): operator weights of its rows, the fictitious-point
// weights, the patched rows after Convert2Tri and the multipliers.  Returns
// whether every patched row is diagonally dominant.
bool block_record(const SetupResult &s, const BetaModel &beta, double theta_dt, int d, int j,
                  int k, const BlockHdr &h, double *rec) {
  const double mt = -theta_dt;
  const Crossing &L = s.cross[d][h.c0];
  const int nrows = h.type == kBlkIrr ? 2 : 3;
  bool dom = true;
  double vd[3][3];
  for (int r = 0; r < nrows; r++) {
    int q[3];
    node_of(d, j, k, L.left_loc + r, q);
    operator_weights(s, beta, d, q[0], q[1], q[2], vd[r]);
  }
  if (h.type == kBlkIrr) {
    for (int r = 0; r < 2; r++)
      for (int m = 0; m < 3; m++) rec[kIrrVd + r * 3 + m] = vd[r][m];
    for (int m = 0; m < 6; m++) {
      rec[kIrrWl + m] = L.wl[m];
      rec[kIrrWr + m] = L.wr[m];
    }
    // patched rows over columns ll-1 .. ll+2 (douglas_adi.cc:405-431)
    double r0[4] = {0, 0, 0, 0}, r1[4] = {0, 0, 0, 0};
    r0[0] = mt * vd[0][0];
    r0[1] = 1 - theta_dt * vd[0][1];
    for (int m = 0; m < 4; m++) r0[m] += mt * vd[0][2] * L.wr[m];
    r1[2] = 1 - theta_dt * vd[1][1];
    r1[3] = mt * vd[1][2];
    for (int m = 0; m < 4; m++) r1[m] += mt * vd[1][0] * L.wl[m];
    // Convert2Tri_irr (douglas_adi.cc:1554-1577)
    double m1 = r1[0] / r0[0];
    for (int m = 0; m < 4; m++) r1[m] -= m1 * r0[m];
    double m2 = r0[3] / r1[3];
    for (int m = 0; m < 4; m++) r0[m] -= m2 * r1[m];
    rec[kIrrLhs + 0] = r0[0];
    rec[kIrrLhs + 1] = r0[1];
    rec[kIrrLhs + 2] = r0[2];
    rec[kIrrLhs + 3] = r1[1];
    rec[kIrrLhs + 4] = r1[2];
    rec[kIrrLhs + 5] = r1[3];
    rec[kIrrMul + 0] = m1;
    rec[kIrrMul + 1] = m2;
    if (!dominant(r0[0], r0[1], r0[2]) || !dominant(r1[1], r1[2], r1[3])) dom = false;
  } else {
    const Crossing &R = s.cross[d][h.c0 + 1];
    for (int r = 0; r < 3; r++)
      for (int m = 0; m < 3; m++) rec[kCorVd + r * 3 + m] = vd[r][m];
    for (int m = 0; m < 9; m++) {
      rec[kCorLwl + m] = L.wl[m];
      rec[kCorLwr + m] = L.wr[m];
      rec[kCorRwl + m] = R.wl[m];
      rec[kCorRwr + m] = R.wr[m];
    }
    // patched rows over columns ll-1 .. ll+3 (douglas_adi.cc:442-482)
    double r0[5] = {0}, r1[5] = {0}, r2[5] = {0};
    r0[0] = mt * vd[0][0];
    r0[1] = 1 - theta_dt * vd[0][1];
    for (int m = 0; m < 5; m++) r0[m] -= theta_dt * vd[0][2] * L.wr[m];
    r1[2] = 1 - theta_dt * vd[1][1];
    for (int m = 0; m < 5; m++) r1[m] -= theta_dt * vd[1][0] * L.wl[m];
    for (int m = 0; m < 5; m++) r1[m] -= theta_dt * vd[1][2] * R.wr[m];
    r2[3] = 1 - theta_dt * vd[2][1];
    r2[4] = mt * vd[2][2];
    for (int m = 0; m < 5; m++) r2[m] -= theta_dt * vd[2][0] * R.wl[m];
    // Convert2Tri_cor (douglas_adi.cc:1637-1675)
    double m10 = r1[0] / r0[0];
    double m20 = r2[0] / r0[0];
    for (int m = 0; m < 5; m++) {
      r1[m] -= m10 * r0[m];
      r2[m] -= m20 * r0[m];
    }
    double m02 = r0[4] / r2[4];
    double m12 = r1[4] / r2[4];
    for (int m = 0; m < 5; m++) {
      r0[m] -= m02 * r2[m];
      r1[m] -= m12 * r2[m];
    }
    double m21 = r2[1] / r1[1];
    for (int m = 0; m < 5; m++) r2[m] -= m21 * r1[m];
    double m01 = r0[3] / r1[3];
    for (int m = 0; m < 5; m++) r0[m] -= m01 * r1[m];
    const double *rows[3] = {r0, r1, r2};
    for (int r = 0; r < 3; r++) {
      rec[kCorLhs + r * 3 + 0] = rows[r][r];
      rec[kCorLhs + r * 3 + 1] = rows[r][r + 1];
      rec[kCorLhs + r * 3 + 2] = rows[r][r + 2];
    }
    rec[kCorMul + 0] = m10;
    rec[kCorMul + 1] = m20;
    rec[kCorMul + 2] = m02;
    rec[kCorMul + 3] = m12;
    rec[kCorMul + 4] = m21;
    rec[kCorMul + 5] = m01;
    for (int r = 0; r < 3; r++)
      if (!dominant(rows[r][r], rows[r][r + 1], rows[r][r + 2])) dom = false;
  }
  return dom;
}
}  // namespace

std::string build_device_table(const SetupResult &s, const BetaModel &beta, double theta_dt,
                               bool with_special, DeviceTableHost &t) {
  const Grid &g = s.grid;
  const int threads = setup_threads();
  std::unordered_map<size_t, int> spec_of;
  t.spec.clear();

  for (int d = 0; d < 3; d++) {
    const int nl = s.n_lines[d];
    const int nj = s.line_nj(d);
    t.blk_ptr[d].assign((size_t)nl + 1, 0);
    t.seq[d].assign((size_t)nl, 0);
    t.n_seq[d] = 0;
    t.blk_hdr[d].clear();
    // block layout in line order: one block per irregular crossing, one per
    // corner pair
    size_t nval = 0;
    for (int lid = 0; lid < nl; lid++) {
      t.blk_ptr[d][lid] = (int)t.blk_hdr[d].size();
      for (int ci = s.line_ptr[d][lid]; ci < s.line_ptr[d][lid + 1];) {
        const bool irr = s.cross[d][ci].id > 0;
        BlockHdr h;
        h.type = irr ? kBlkIrr : kBlkCor;
        h.ll = s.cross[d][ci].left_loc;
        h.c0 = ci;
        h.off = (int)nval;
        nval += irr ? kBlkIrrLen : kBlkCorLen;
        t.blk_hdr[d].push_back(h);
        ci += irr ? 1 : 2;
      }
    }
    t.blk_ptr[d][nl] = (int)t.blk_hdr[d].size();
    if (nval > (size_t)INT32_MAX) return "MIB table too large";
    t.blk_val[d].assign(nval, 0.0);
    t.cr_val[d].assign(s.cross[d].size() * kCrLen, 0.0);
    t.cls0[d].assign((size_t)nl, 0);
    // records, jump records, first-node class per line and the check that
    // the classification of every node is that class flipped at each
    // crossing to its left (the kernels recover it that way): lines in
    // parallel, each writing only its own entries
    const int n = g.n[d];
    const int chunks = chunk_count(nl, threads);
    std::vector<uint8_t> bad((size_t)chunks, 0);
    parallel_chunks(nl, chunks, threads, [&](int ch, int l0, int l1) {
      for (int lid = l0; lid < l1; lid++) {
        const int j = lid % nj, k = lid / nj;
        const int c0 = s.line_ptr[d][lid], c1 = s.line_ptr[d][lid + 1];
        for (int ci = c0; ci < c1; ci++)
          jump_record(s.cross[d][ci], &t.cr_val[d][(size_t)ci * kCrLen]);
        for (int bi = t.blk_ptr[d][lid]; bi < t.blk_ptr[d][lid + 1]; bi++) {
          const BlockHdr &h = t.blk_hdr[d][bi];
          if (!block_record(s, beta, theta_dt, d, j, k, h, &t.blk_val[d][(size_t)h.off]))
            t.seq[d][lid] = 1;
        }
        int q[3];
        node_of(d, j, k, 0, q);
        bool out = s.ind[g.idx(q[0], q[1], q[2])] > 0;
        t.cls0[d][lid] = out ? 1 : 0;
        int c = c0;
        for (int i = 0; i < n; i++) {
          while (c < c1 && s.cross[d][c].left_loc < i) {
            out = !out;
            c++;
          }
          node_of(d, j, k, i, q);
          if ((s.ind[g.idx(q[0], q[1], q[2])] > 0) != out) {
            bad[(size_t)ch] = 1;
            return;
          }
        }
      }
    });
    for (uint8_t b : bad)
      if (b) return "node classification is not the line's first class flipped at its crossings";
    for (int lid = 0; lid < nl; lid++) t.n_seq[d] += t.seq[d][lid];
    // Douglas stage 1: the nodes on patched rows, each with its block and
    // row per direction, numbered in the order first met
    if (with_special) {
      for (int lid = 0; lid < nl; lid++) {
        const int j = lid % nj, k = lid / nj;
        for (int bidx = t.blk_ptr[d][lid]; bidx < t.blk_ptr[d][lid + 1]; bidx++) {
          const BlockHdr &h = t.blk_hdr[d][bidx];
          const int nrows = h.type == kBlkIrr ? 2 : 3;
          for (int r = 0; r < nrows; r++) {
            int q[3];
            node_of(d, j, k, h.ll + r, q);
            size_t node = g.idx(q[0], q[1], q[2]);
            auto it = spec_of.find(node);
            int sidx;
            if (it == spec_of.end()) {
              sidx = (int)(t.spec.size() / kSpecLen);
              spec_of.emplace(node, sidx);
              int recs[kSpecLen] = {q[0], q[1], q[2], -1, -1, -1, -1, -1, -1, 0};
              t.spec.insert(t.spec.end(), recs, recs + kSpecLen);
            } else {
              sidx = it->second;
            }
            t.spec[(size_t)sidx * kSpecLen + 3 + d] = bidx;
            t.spec[(size_t)sidx * kSpecLen + 6 + d] = r;
          }
        }
      }
    }
    t.n_blocks[d] = (int)t.blk_hdr[d].size();
  }

  // axis tables
  for (int a = 0; a < 3; a++) {
    int n = g.n[a];
    t.tab[a].assign((size_t)kTabCount * n, 0.0);
    for (int i = 0; i < n; i++) {
      double x = g.c[a][i], h = g.h[a];
      t.tab[a][(size_t)kTabSin * n + i] = std::sin(x);
      t.tab[a][(size_t)kTabCos * n + i] = std::cos(x);
      t.tab[a][(size_t)kTabSinM * n + i] = std::sin(x - h / 2);
      t.tab[a][(size_t)kTabSinP * n + i] = std::sin(x + h / 2);
      t.tab[a][(size_t)kTabCosM * n + i] = std::cos(x - h / 2);
      t.tab[a][(size_t)kTabCosP * n + i] = std::cos(x + h / 2);
    }
  }

  PhysConst &p = t.phys;
  std::memset(&p, 0, sizeof(p));
  p.beta_kind = beta.kind;
  p.b_in = beta.p[0];
  p.b_out = beta.p[1];
  p.f_in = 3.0 * beta.p[0] - 1.0;
  p.f_out = 3.0 * beta.p[1] - 1.0;
  for (int a = 0; a < 3; a++) {
    double h = g.h[a];
    p.hh[a] = h * h;
    p.ihh[a] = 1.0 / (h * h);
    double bs[2] = {beta.p[0], beta.p[1]};
    for (int side = 0; side < 2; side++) {
      double b = bs[side];
      p.vdc[a][side][0] = b / (h * h);
      p.vdc[a][side][1] = -(b + b) / (h * h);
      p.vdc[a][side][2] = b / (h * h);
    }
  }
  return std::string();
}

}  // namespace mts
