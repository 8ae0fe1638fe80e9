// The host setup gives the same result on any number of threads: node
// classification, crossings with their MIB weights, the device table and the
// stage records of every stage kind, compared bit for bit between one thread
// and several, and the same error message where the setup refuses a
// geometry.  Built with ThreadSanitizer, so a data race between the chunks
// also fails the check.
//
//   setup_threads_check <nodes per axis> <mib 1|2> <surface ellipsoid|torus|bumpy>
//                       <p0> <p1> [p2 p3] <beta const|sincos> <b0> <b1> <threads>...
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "device_types.hpp"
#include "host_parallel.hpp"
#include "host_setup.hpp"
#include "host_table.hpp"
#include "stage_records.hpp"

using namespace mts;

namespace {

struct Built {
  std::string err;
  SetupResult res;
  DeviceTableHost t[2];                  // Douglas (with special nodes), theta = 1/2
  std::vector<std::vector<double>> rec;  // every stage kind, both tables
};

const int kKinds[] = {kRhsCopy | kFlagSolve, kRhsDouglas | kFlagSolve,
                      kRhsCopy | kFlagSolve | kFlagLhsJump, kRhsCN | kFlagSolve | kFlagLhsJump,
                      kRhsCN, kRhsTSZ | kFlagSolve | kFlagLhsJump};

void build(const Grid &g, const Surface &s, const BetaModel &beta, int mib, int threads,
           Built &b) {
  set_setup_thread_share(threads);
  BuildOptions opt;
  opt.mib = mib;
  b.err = build_setup(g, s, beta, opt, b.res);
  if (!b.err.empty()) return;
  const double dt = 2.0 / (g.n[0] - 1);
  for (int v = 0; v < 2; v++) {
    const double tdt = v == 0 ? dt : 0.5 * dt;
    b.err = build_device_table(b.res, beta, tdt, v == 0, b.t[v]);
    if (!b.err.empty()) return;
    for (int d = 0; d < 3; d++)
      for (int kind : kKinds) {
        std::vector<double> r;
        b.err = build_stage_records(b.res, b.t[v], d, kind, tdt, dt, 0.5 * dt, 0.5 * dt, r);
        if (!b.err.empty()) return;
        b.rec.push_back(std::move(r));
      }
  }
}

template <class T>
bool same(const std::vector<T> &a, const std::vector<T> &b) {
  return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

bool same_crossing(const Crossing &a, const Crossing &b) {
  return a.dir == b.dir && a.j == b.j && a.k == b.k && a.id == b.id &&
         a.left_loc == b.left_loc && a.nw == b.nw &&
         std::memcmp(&a.gamma, &b.gamma, sizeof a.gamma) == 0 &&
         std::memcmp(a.coord, b.coord, sizeof a.coord) == 0 &&
         std::memcmp(a.normal, b.normal, sizeof a.normal) == 0 &&
         std::memcmp(a.p, b.p, sizeof a.p) == 0 &&
         std::memcmp(&a.pinv_d0, &b.pinv_d0, sizeof a.pinv_d0) == 0 &&
         std::memcmp(&a.beta_in, &b.beta_in, sizeof a.beta_in) == 0 &&
         std::memcmp(&a.beta_out, &b.beta_out, sizeof a.beta_out) == 0 &&
         std::memcmp(a.wl, b.wl, sizeof a.wl) == 0 && std::memcmp(a.wr, b.wr, sizeof a.wr) == 0;
}

// "" when equal, else what differs
std::string compare(const Built &a, const Built &b) {
  if (a.err != b.err) return "error message: '" + a.err + "' vs '" + b.err + "'";
  if (!a.err.empty()) return "";
  if (!same(a.res.ind, b.res.ind)) return "classification";
  if (a.res.singular_warnings != b.res.singular_warnings) return "singular warnings";
  for (int d = 0; d < 3; d++) {
    if (a.res.cross[d].size() != b.res.cross[d].size()) return "crossing count";
    for (size_t i = 0; i < a.res.cross[d].size(); i++)
      if (!same_crossing(a.res.cross[d][i], b.res.cross[d][i])) return "crossing record";
    if (!same(a.res.line_ptr[d], b.res.line_ptr[d])) return "line_ptr";
    if (a.res.corner_pairs[d] != b.res.corner_pairs[d]) return "corner pairs";
  }
  for (int v = 0; v < 2; v++) {
    const DeviceTableHost &x = a.t[v], &y = b.t[v];
    for (int d = 0; d < 3; d++) {
      if (!same(x.blk_ptr[d], y.blk_ptr[d])) return "blk_ptr";
      if (x.blk_hdr[d].size() != y.blk_hdr[d].size()) return "block count";
      for (size_t i = 0; i < x.blk_hdr[d].size(); i++) {
        const BlockHdr &p = x.blk_hdr[d][i], &q = y.blk_hdr[d][i];
        if (p.type != q.type || p.ll != q.ll || p.c0 != q.c0 || p.off != q.off) return "block header";
      }
      if (!same(x.blk_val[d], y.blk_val[d])) return "block records";
      if (!same(x.cr_val[d], y.cr_val[d])) return "jump records";
      if (!same(x.seq[d], y.seq[d]) || x.n_seq[d] != y.n_seq[d]) return "sequential-line flags";
      if (!same(x.cls0[d], y.cls0[d])) return "first-node classes";
      if (!same(x.tab[d], y.tab[d])) return "axis tables";
      if (x.n_blocks[d] != y.n_blocks[d]) return "n_blocks";
    }
    if (!same(x.spec, y.spec)) return "special nodes";
  }
  if (a.rec.size() != b.rec.size()) return "stage record count";
  for (size_t i = 0; i < a.rec.size(); i++)
    if (!same(a.rec[i], b.rec[i])) return "stage records";
  return "";
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 9) {
    std::fprintf(stderr, "usage: setup_threads_check nodes mib surface p.. beta b0 b1 threads..\n");
    return 2;
  }
  int nn[3];
  if (std::sscanf(argv[1], "%dx%dx%d", &nn[0], &nn[1], &nn[2]) != 3)
    nn[0] = nn[1] = nn[2] = std::atoi(argv[1]);
  const int mib = std::atoi(argv[2]);
  const std::string sk = argv[3];
  int ai;
  std::unique_ptr<Surface> surf;
  if (sk == "ellipsoid") {
    surf.reset(new SurfaceEllipsoid(atof(argv[4]), atof(argv[5]), atof(argv[6]), atof(argv[7])));
    ai = 8;
  } else if (sk == "torus") {
    surf.reset(new SurfaceTorus(atof(argv[4]), atof(argv[5])));
    ai = 6;
  } else if (sk == "bumpy") {
    surf.reset(new SurfaceBumpy(atof(argv[4]), atof(argv[5]), atof(argv[6]), atof(argv[7])));
    ai = 8;
  } else {
    return 2;
  }
  if (argc < ai + 4) return 2;
  BetaModel beta;
  beta.kind = std::string(argv[ai]) == "sincos" ? 1 : 0;
  beta.p[0] = atof(argv[ai + 1]);
  beta.p[1] = atof(argv[ai + 2]);
  double lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};
  const Grid g = make_grid(lo, hi, nn);
  Built one;
  build(g, *surf, beta, mib, 1, one);
  size_t ncross = 0;
  for (int d = 0; d < 3; d++) ncross += one.res.cross[d].size();
  for (int i = ai + 3; i < argc; i++) {
    const int th = std::atoi(argv[i]);
    Built many;
    build(g, *surf, beta, mib, th, many);
    const std::string diff = compare(one, many);
    if (!diff.empty()) {
      std::printf("differs on %d threads: %s\n", th, diff.c_str());
      return 1;
    }
  }
  std::printf("ok: %zu crossings%s%s\n", ncross, one.err.empty() ? "" : ", refused: ",
              one.err.c_str());
  return 0;
}
