// Stage records: each MIB block's patched rows as linear forms in the
// block's input rows and time factors, for one stage kind (device_types.hpp,
// kRec*).  Derived from block_rhs, the reference-order computation of the
// patched rows (douglas_adi.cc:116-195, lod.cc:414-440 and :887-1078), by
// evaluating it on unit inputs: block_rhs is linear in the input rows and in
// e_jump, e_f1 and e_f2, so each unit input yields one coefficient column.
#include "stage_records.hpp"

#include <cstring>

#include "device_core.hpp"
#include "host_parallel.hpp"

namespace mts {

namespace {

template <int DIR, int FLAGS>
void fill(const SetupResult &s, const DeviceTableHost &t, const LineArgs &p0,
          std::vector<double> &rec) {
  constexpr bool HAS_S = (FLAGS & 3) == kRhsDouglas;
  const int d = DIR;
  const int n = s.grid.n[d];
  const int nl = s.n_lines[d];
  const int nj = s.line_nj(d);
  // lines in parallel: a block's record depends on its own line alone
  const int threads = setup_threads();
  parallel_chunks(nl, chunk_count(nl, threads), threads, [&](int, int l0, int l1) {
    LineArgs p = p0;
    // unit inputs: one slot set for an evaluation and cleared after it
    std::vector<double> la((size_t)p.q.lds_len, 0.0), lb((size_t)p.q.lds_len, 0.0);
    for (int lid = l0; lid < l1; lid++) {
      const LineMeta lm = line_meta(p.mib, lid, true);
      const int j = lid % nj, k = lid / nj;
      for (int bi = lm.bb; bi < lm.be; bi++) {
        const BlockHdr h = t.blk_hdr[d][bi];
        const int nr = h.type == kBlkIrr ? 2 : 3;
        double *out = &rec[(size_t)bi * kRecLen];
        auto eval = [&](int col, int which, double ej, double e1, double e2) {
          double *unit =
              col >= 0 ? &(which == 0 ? la : lb)[lds_slot(p.q, h.ll - 1 + col)] : nullptr;
          if (unit) *unit = 1.0;
          p.e_jump = ej;
          p.e_f1 = e1;
          p.e_f2 = e2;
          const BlockRhs v = block_rhs<DIR, FLAGS>(p, j, k, lm, h, la.data(), lb.data());
          if (unit) *unit = 0.0;
          return v;
        };
        auto put = [&](const BlockRhs &v, int off) {
          const double vv[3] = {v.v0, v.v1, v.v2};
          for (int r = 0; r < nr; r++) out[r * kRecRow + off] = vv[r];
        };
        for (int c = 0; c < 5; c++) {
          if (h.ll - 1 + c > n - 1) continue;
          put(eval(c, 0, 0.0, 0.0, 0.0), kRecWA + c);
          if (HAS_S) put(eval(c, 1, 0.0, 0.0, 0.0), kRecWS + c);
        }
        put(eval(-1, 0, 1.0, 0.0, 0.0), kRecK);
        if ((FLAGS & 3) == kRhsTSZ) {
          put(eval(-1, 0, 0.0, 1.0, 0.0), kRecK1);
          put(eval(-1, 0, 0.0, 0.0, 1.0), kRecK2);
        }
        const double *v = &t.blk_val[d][(size_t)h.off];
        const double *lhs = v + (h.type == kBlkIrr ? kIrrLhs : kCorLhs);
        for (int r = 0; r < nr; r++)
          for (int q = 0; q < 3; q++) out[r * kRecRow + kRecLhs + q] = lhs[r * 3 + q];
      }
    }
  });
}

template <int DIR>
bool dispatch(int flags, const SetupResult &s, const DeviceTableHost &t, const LineArgs &p,
              std::vector<double> &rec) {
  switch (flags & ~(kFlagEpiF | kFlagSeq | kFlagPre1)) {
    case kRhsCopy | kFlagSolve: fill<DIR, kRhsCopy | kFlagSolve>(s, t, p, rec); return true;
    case kRhsCopy | kFlagSolve | kFlagLhsJump:
      fill<DIR, kRhsCopy | kFlagSolve | kFlagLhsJump>(s, t, p, rec);
      return true;
    case kRhsDouglas | kFlagSolve: fill<DIR, kRhsDouglas | kFlagSolve>(s, t, p, rec); return true;
    case kRhsCN | kFlagSolve | kFlagLhsJump:
      fill<DIR, kRhsCN | kFlagSolve | kFlagLhsJump>(s, t, p, rec);
      return true;
    case kRhsCN: fill<DIR, kRhsCN>(s, t, p, rec); return true;
    case kRhsTSZ | kFlagSolve | kFlagLhsJump:
      fill<DIR, kRhsTSZ | kFlagSolve | kFlagLhsJump>(s, t, p, rec);
      return true;
    default: return false;
  }
}

}  // namespace

std::string build_stage_records(const SetupResult &s, const DeviceTableHost &t, int d, int flags,
                                double theta_dt, double dt, double c_f1, double c_f2,
                                std::vector<double> &rec) {
  LineArgs p;
  std::memset(&p, 0, sizeof(p));
  for (int a = 0; a < 3; a++) {
    p.g.n[a] = s.grid.n[a];
    p.g.tab[a] = t.tab[a].data();
  }
  p.mib.blk_ptr = t.blk_ptr[d].data();
  p.mib.hdr = t.blk_hdr[d].data();
  p.mib.val = t.blk_val[d].data();
  p.mib.cr = t.cr_val[d].data();
  p.mib.cls0 = t.cls0[d].data();
  p.P = t.phys;
  p.theta_dt = theta_dt;
  p.dt = dt;
  p.c_f1 = c_f1;
  p.c_f2 = c_f2;
  if (!make_line_part(s.grid.n[d], p.q)) return "line length not supported";
  rec.assign((size_t)t.n_blocks[d] * kRecLen, 0.0);
  bool ok = d == 0 ? dispatch<0>(flags, s, t, p, rec)
                   : d == 1 ? dispatch<1>(flags, s, t, p, rec) : dispatch<2>(flags, s, t, p, rec);
  if (!ok) return "unknown stage kind";
  return std::string();
}

}  // namespace mts
