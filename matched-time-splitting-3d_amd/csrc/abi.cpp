// C ABI of the Matched-TS per-step path (include/mts.h).
#include "mts.h"
#include "ref_checks.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "abi_internal.hpp"
#include "device_core.hpp"
#include "diagnostics.hpp"
#include "host_setup.hpp"
#include "host_table.hpp"
#include "kernels.hpp"
#include "regular_solve.hpp"
#include "stage_records.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

}  // namespace

int mts::set_error(int code, const char *msg) { return fail(code, msg ? msg : ""); }

namespace {

#define MTS_HIP(call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail(MTS_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)

std::unique_ptr<mts::Surface> make_surface(const mts_surface_desc *sd, std::string &err) {
  switch (sd->kind) {
    case MTS_SURF_ELLIPSOID:
      return std::unique_ptr<mts::Surface>(
          new mts::SurfaceEllipsoid(sd->p[0], sd->p[1], sd->p[2], sd->p[3]));
    case MTS_SURF_TORUS:
      return std::unique_ptr<mts::Surface>(new mts::SurfaceTorus(sd->p[0], sd->p[1]));
    case MTS_SURF_BUMPY:
      return std::unique_ptr<mts::Surface>(
          new mts::SurfaceBumpy(sd->p[0], sd->p[1], sd->p[2], sd->p[3]));
    case MTS_SURF_CALLBACK:
      if (!sd->phi || !sd->grad) {
        err = "callback surface needs phi and grad";
        return nullptr;
      }
      return std::unique_ptr<mts::Surface>(new mts::SurfaceCallback(sd->phi, sd->grad, sd->user));
    default:
      err = "unknown surface kind";
      return nullptr;
  }
}

bool make_beta(const mts_physics_desc *pd, mts::BetaModel &b, std::string &err) {
  if (pd->beta_kind != MTS_BETA_CONST && pd->beta_kind != MTS_BETA_SINCOS) {
    err = "unknown beta kind";
    return false;
  }
  b.kind = pd->beta_kind;
  for (int i = 0; i < 4; i++) b.p[i] = pd->bp[i];
  return true;
}

// Device buffer.  With a stream it is allocated and freed stream-ordered
// (hipMallocAsync / hipFreeAsync from the device's memory pool), so creating
// and destroying instances does not synchronize the device while other
// instances' steps are in flight; the stream must outlive the buffer.
template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  hipStream_t st = nullptr;
  ~DevBuf() {
    if (p) (void)(st ? hipFreeAsync(p, st) : hipFree(p));
  }
  hipError_t alloc(size_t count, hipStream_t s = nullptr) {
    n = count;
    st = s;
    if (count == 0) return hipSuccess;
    if (s) return hipMallocAsync(reinterpret_cast<void **>(&p), count * sizeof(T), s);
    return hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
  }
  // Allocates and copies on stream s.  The copy is asynchronous: h must stay
  // valid, and no other stream may read the buffer, until s is synchronized.
  size_t bytes() const { return p ? n * sizeof(T) : 0; }
  hipError_t upload(const T *h, size_t count, hipStream_t s) {
    hipError_t e = alloc(count, s);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

// Owns a stream; declared before an instance's buffers so that it is
// destroyed after them (their stream-ordered frees need it).
struct StreamHolder {
  hipStream_t s = nullptr;
  ~StreamHolder() {
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }
};

}  // namespace

struct mts_setup {
  mts::SetupResult res;
  mts::BetaModel beta;
};

struct mts_lu {
  std::vector<double> a;  // the decomposed matrix, kept for mts_lu_mprove
  mts::DenseLU lu;
  mts_lu(int n, const double *m) : a(m, m + (size_t)n * n), lu(n, m) {}
};

struct mts_ctx {
  int device = 0;
  hipDeviceProp_t prop;
  void *norms_scratch = nullptr;  // device address of the library's fixed norms array
  // mts_tune_lanes' choices: (nx, ny, nz, method, precision) -> split
  std::mutex tune_mu;
  std::map<std::array<int, 5>, mts::Lanes> tuned;
};

bool mts::ctx_tuned_lanes(mts_ctx *ctx, const int n[3], int method, int precision,
                          mts::Lanes &out) {
  if (!ctx) return false;
  std::lock_guard<std::mutex> lk(ctx->tune_mu);
  auto it = ctx->tuned.find({n[0], n[1], n[2], method, precision});
  if (it == ctx->tuned.end()) return false;
  out = it->second;
  return true;
}

struct mts_instance {
  mts_ctx *ctx = nullptr;
  int method = 0;
  int min_lanes = 16;  // lane-group lower bound of the line split (tuning)
  int gen_lanes = 16;  // lanes per line of the general tiles' solve (tuning)
  bool douglas_fused = true;  // Douglas stage 1 inside the x sweep (kFlagPre1)
  double dt = 0.0, theta_dt = 0.0;
  mts::Grid grid;
  int px = 0;
  size_t field_elems = 0;
  StreamHolder sh;                // the instance's stream (outlives the buffers below)
  hipStream_t stream = nullptr;   // == sh.s
  StreamHolder side_sh;           // second stream of the split solve stages
  mts::SideStream side;           // side_sh.s and its fork / join events (null: one stream)
  int precision = MTS_FP64;       // field storage and line solve
  DevBuf<double> u, r;            // solution (FP64 form); Douglas stage scratch
  DevBuf<float> uf, rf;           // the same in the FP32 form
  bool f32() const { return precision == MTS_FP32; }
  void *field() { return f32() ? static_cast<void *>(uf.p) : static_cast<void *>(u.p); }
  void *stage() { return f32() ? static_cast<void *>(rf.p) : static_cast<void *>(r.p); }
  DevBuf<int8_t> ind;
  DevBuf<double> tab[3];
  DevBuf<int> blk_ptr[3];
  DevBuf<mts::BlockHdr> blk_hdr[3];
  DevBuf<double> blk_val[3];
  DevBuf<double> cr_val[3];
  DevBuf<uint8_t> seq[3];
  DevBuf<uint8_t> cls0[3];
  // stage records per (direction, stage kind)
  std::map<int, std::unique_ptr<DevBuf<double>>> recs;
  static int rec_key(int d, int flags) {
    return d * 64 + (flags & ~(mts::kFlagEpiF | mts::kFlagSeq | mts::kFlagPre1));
  }
  const double *rec_for(int d, int flags) const {
    auto it = recs.find(rec_key(d, flags));
    return it == recs.end() ? nullptr : it->second->p;
  }
  // regular lines' solve tables per direction: [0] guarded, [1] unguarded
  DevBuf<uint8_t> regt[3][2];
  int reg_side[3] = {0, 0, 0};
  DevBuf<int> tiles[3][2];  // per direction: regular tiles, the others (tile_lists)
  DevBuf<int> spec;
  int nspec = 0;
  DevBuf<double> part_max, part_sq;
  DevBuf<double> diag;  // interface diagnostics: el, er, je per crossing, then 4 sums (on first use)
  DevBuf<long long> part_idx;
  mts::PhysConst phys;
  mts::DevGeom geom;
  mts::DevMib mib[3];
  // per-stage event timing
  bool profile = false;
  // Recording a step instead of launching it (mts_batch_run_multi): solve
  // line stages as their arguments, every other launch as a call on a given
  // stream.
  struct StepOp {
    int line = 0;  // 1: a split solve stage (dir, flags, a); 2: the faces (fa);
                   // 3: Douglas stage 1 (r1, sp, pairs_too); 0: fn
    int dir = 0, flags = 0;
    mts::LineArgs a;
    mts::FaceArgs fa;
    mts::Rhs1Args r1;
    mts::SpecArgs sp;
    bool pairs_too = false;
    std::function<hipError_t(hipStream_t)> fn;
  };
  std::vector<StepOp> *rec = nullptr;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, size_t>> pending;  // stage id, first event index
  std::vector<std::string> stage_names;
  std::map<std::string, int> stage_id;
  std::vector<double> stage_ms;
  std::vector<int> stage_n;
  // the time loop captured as one graph (mts_run_graph), kept for repeated runs
  hipGraphExec_t gexec = nullptr;
  double g_t0 = 0.0, g_tf = -1.0, g_tend = 0.0;
  int g_steps = 0;
  ~mts_instance() {
    if (stream) (void)hipStreamSynchronize(stream);  // pending copies may read host staging
    if (side.s) (void)hipStreamSynchronize(side.s);
    if (side.fork) (void)hipEventDestroy(side.fork);
    if (side.join) (void)hipEventDestroy(side.join);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
  }
  size_t device_bytes() const {
    size_t b = u.bytes() + uf.bytes() + r.bytes() + rf.bytes() + ind.bytes() + spec.bytes() + part_max.bytes() +
               part_sq.bytes() + part_idx.bytes();
    for (int d = 0; d < 3; d++)
      b += tab[d].bytes() + blk_ptr[d].bytes() + blk_hdr[d].bytes() + blk_val[d].bytes() +
           cr_val[d].bytes() + seq[d].bytes() + cls0[d].bytes() + regt[d][0].bytes() +
           regt[d][1].bytes() + tiles[d][0].bytes() + tiles[d][1].bytes();
    for (auto &kv : recs) b += kv.second->bytes();
    return b;
  }
  int stage(const char *name) {
    auto it = stage_id.find(name);
    if (it != stage_id.end()) return it->second;
    int id = (int)stage_names.size();
    stage_id[name] = id;
    stage_names.push_back(name);
    stage_ms.push_back(0.0);
    stage_n.push_back(0);
    return id;
  }
  hipEvent_t event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  hipError_t drain() {
    if (pending.empty()) return hipSuccess;
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return e;
    for (auto &p : pending) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, ev_pool[p.second], ev_pool[p.second + 1]);
      if (e != hipSuccess) return e;
      stage_ms[p.first] += ms;
      stage_n[p.first] += 1;
    }
    pending.clear();
    ev_used = 0;
    return hipSuccess;
  }
};

namespace {
// Runs one launch, bracketed by events when profiling is on.
template <class F>
hipError_t timed(mts_instance *in, const char *name, F launch) {
  if (!in->profile) return launch();
  if (in->ev_used + 2 > 4096) {
    hipError_t e = in->drain();
    if (e != hipSuccess) return e;
  }
  int id = in->stage(name);
  size_t first = in->ev_used;
  hipEvent_t a = in->event(), b = in->event();
  if (!a || !b) return hipErrorOutOfMemory;
  hipError_t e = hipEventRecord(a, in->stream);
  if (e != hipSuccess) return e;
  e = launch();
  if (e != hipSuccess) return e;
  e = hipEventRecord(b, in->stream);
  if (e != hipSuccess) return e;
  in->pending.push_back({id, first});
  return hipSuccess;
}
}  // namespace

// Host half of instance creation: everything but device memory, so that a
// serving loop can build it on a setup thread and know the exact device
// footprint before admitting the solve.
struct mts::InstancePlan {
  int method = 0;
  int precision = MTS_FP64;
  int min_lanes = 16;
  int gen_lanes = 16;  // lanes per line of the general tiles' solve (tuning)
  double dt = 0.0, theta_dt = 0.0;
  mts::Grid grid;
  int px = 0;
  size_t field_elems = 0;
  mts::DeviceTableHost t;
  std::vector<int8_t> ind;                              // x-padded node classification
  std::vector<std::pair<int, std::vector<double>>> recs;  // (record key, stage records)
  std::vector<uint8_t> reg[3][2];  // regular lines' solve tables (field type), guarded / unguarded
  int reg_side[3] = {0, 0, 0};     // elements per side
  std::vector<int> tiles[3][2];    // per direction: regular tiles, the others
  size_t device_bytes = 0;
};

namespace {
// The regular lines' solve tables of direction d for an instance: the
// outside side, then, when beta is constant inside, the inside side (lines
// inside a variable beta are never regular), each reg_side_len elements of
// the field type R, for the guarded solve and, when the split lets the
// kernels use it, the unguarded one.  The rows' matrix coefficients are
// formed exactly as the kernels form regular rows (in R).
template <typename R>
bool build_regular_tables(const mts::PhysConst &P, const mts::LinePart &q, double theta_dt, int d,
                          std::vector<uint8_t> out[2], int &side_len) {
  side_len = mts::reg_side_len(q.L, q.mmax);
  for (int full = 0; full < 2; full++) {
    out[full].clear();
    if (full && q.m != q.mmax) continue;
    std::vector<double> tab;
    for (int side = 1; side >= 0; side--) {
      if (side == 0 && P.beta_kind != 0) continue;
      const R rt = (R)theta_dt;
      const R w0 = (R)P.vdc[d][side][0], w1 = (R)P.vdc[d][side][1], w2 = (R)P.vdc[d][side][2];
      const R a = -rt * w0, b = R(1) - rt * w1, c = -rt * w2;
      if (!mts::regular_factors(q, q.mmax, full != 0, (double)a, (double)b, (double)c, tab))
        return false;
    }
    std::vector<R> tr(tab.begin(), tab.end());
    // whole 16-byte units (the kernels stage the tables with 16-byte loads)
    out[full].assign((tr.size() * sizeof(R) + 15) / 16 * 16, 0);
    std::memcpy(out[full].data(), tr.data(), tr.size() * sizeof(R));
  }
  return true;
}
}  // namespace

void mts::PlanDeleter::operator()(mts::InstancePlan *p) const { delete p; }

size_t mts::plan_device_bytes(const mts::InstancePlan &p) { return p.device_bytes; }

int mts::plan_instance(const mts_setup *setup, const mts_physics_desc *physics, int method,
                       double dt, int precision, mts::PlanPtr &out, const mts::Lanes *lanes) {
  if (!setup || !physics) return fail(MTS_E_ARG, "null argument");
  if (method < MTS_ADI || method > MTS_TS) return fail(MTS_E_ARG, "unknown method");
  if (precision != MTS_FP64 && precision != MTS_FP32) return fail(MTS_E_ARG, "unknown precision");
  if (!(dt > 0.0)) return fail(MTS_E_ARG, "dt must be positive");
  mts::BetaModel beta;
  std::string err;
  if (!make_beta(physics, beta, err)) return fail(MTS_E_ARG, err);
  if (beta.kind != setup->beta.kind || beta.p[0] != setup->beta.p[0] ||
      beta.p[1] != setup->beta.p[1])
    return fail(MTS_E_ARG, "physics differs from the one the setup was built with");
  if (method == MTS_ADI) {
    const int *cp = setup->res.corner_pairs;
    if (cp[0] + cp[1] + cp[2] > 0)
      return fail(MTS_E_DISABLED,
                  "Douglas-ADI on an interface with corner pairs is switched off on the GPU: "
                  "its step raised a GPU memory aperture violation (dispatch reported as "
                  "spec_kernel) whose cause is not found; Douglas-ADI without corner pairs, "
                  "and LOD-IE, LOD-CN and TS with or without them, are not affected");
  }
  try {
    mts::PlanPtr p(new mts::InstancePlan());
    p->method = method;
    p->precision = precision;
    p->dt = dt;
    p->theta_dt = (method == MTS_ADI || method == MTS_LOD_IE) ? dt : 0.5 * dt;
    p->grid = setup->res.grid;
    const mts::Grid &g = p->grid;
    if (lanes && lanes->min_lanes) {
      // a chosen split (mts_tune_lanes)
      p->min_lanes = lanes->min_lanes;
      p->gen_lanes = lanes->gen_lanes;
    } else {
      // MTS_LINE_LANES=32|64 raises the lane-group lower bound of the line
      // split (every split computes the same solution to rounding)
      if (const char *ev = std::getenv("MTS_LINE_LANES")) {
        const int v = std::atoi(ev);
        if (v == 16 || v == 32 || v == 64) p->min_lanes = v;
      }
      // MTS_GEN_LANES=32|64: the same for the general tiles of the solve stages
      if (const char *ev = std::getenv("MTS_GEN_LANES")) {
        const int v = std::atoi(ev);
        if (v == 16 || v == 32 || v == 64) p->gen_lanes = v;
      }
    }
    for (int d = 0; d < 3; d++) {
      mts::LinePart q;
      if (!mts::make_line_part(g.n[d], q, p->min_lanes))
        return fail(MTS_E_ARG, "line length above 1024 nodes is not supported");
    }
    // x rows start on 128-byte lines: 16 doubles, 32 floats
    const int align = precision == MTS_FP32 ? 32 : 16;
    p->px = (g.n[0] + align - 1) / align * align;
    p->field_elems = (size_t)p->px * g.n[1] * g.n[2];
    mts::DeviceTableHost &t = p->t;
    const std::string terr =
        mts::build_device_table(setup->res, beta, p->theta_dt, method == MTS_ADI, t);
    if (!terr.empty()) return fail(MTS_E_GEOMETRY, terr);
    p->ind.assign(p->field_elems, 1);
    for (int iz = 0; iz < g.n[2]; iz++)
      for (int iy = 0; iy < g.n[1]; iy++)
        std::memcpy(&p->ind[((size_t)iz * g.n[1] + iy) * p->px],
                    &setup->res.ind[g.idx(0, iy, iz)], g.n[0]);
    // Douglas-ADI: the nodes spec_node writes carry +-2 (the sign is still
    // the side), so the regular stage-1 stencil leaves them alone
    for (size_t r = 0; r + mts::kSpecLen <= t.spec.size(); r += mts::kSpecLen) {
      int8_t &v = p->ind[((size_t)t.spec[r + 2] * g.n[1] + t.spec[r + 1]) * p->px + t.spec[r]];
      v = v > 0 ? 2 : -2;
    }
    using namespace mts;
    std::vector<std::pair<int, int>> kinds;  // (direction, stage kind) this method launches
    for (int d = 0; d < 3; d++) {
      if (method == MTS_ADI) kinds.push_back({d, d == 0 ? kRhsCopy | kFlagSolve : kRhsDouglas | kFlagSolve});
      else if (method == MTS_LOD_IE) kinds.push_back({d, kRhsCopy | kFlagSolve | kFlagLhsJump});
      else if (method == MTS_LOD_CN) kinds.push_back({d, kRhsCN | kFlagSolve | kFlagLhsJump});
      else if (d < 2) {
        kinds.push_back({d, kRhsCN});
        kinds.push_back({d, kRhsCopy | kFlagSolve | kFlagLhsJump});
      } else {
        kinds.push_back({d, kRhsTSZ | kFlagSolve | kFlagLhsJump});
      }
    }
    for (auto &kd : kinds) {
      std::vector<double> rec;
      const std::string rerr = build_stage_records(setup->res, t, kd.first, kd.second,
                                                   p->theta_dt, dt, 0.5 * dt, 0.5 * dt, rec);
      if (!rerr.empty()) return fail(MTS_E_ARG, rerr);
      p->recs.push_back({mts_instance::rec_key(kd.first, kd.second), std::move(rec)});
    }
    for (int d = 0; d < 3; d++) {
      mts::LinePart q;
      mts::make_line_part(g.n[d], q, p->min_lanes);
      const bool ok = precision == MTS_FP32
                          ? build_regular_tables<float>(t.phys, q, p->theta_dt, d, p->reg[d], p->reg_side[d])
                          : build_regular_tables<double>(t.phys, q, p->theta_dt, d, p->reg[d], p->reg_side[d]);
      if (!ok) return fail(MTS_E_ARG, "zero pivot in the regular lines' line solve");
      const int nj = d == 0 ? g.n[1] : g.n[0];
      const int nk = d == 2 ? g.n[1] : g.n[2];
      mts::stage_tiles(t.blk_ptr[d].data(), t.cls0[d].data(), t.phys.beta_kind == 0, d, g.n[d],
                       nj, nk, p->min_lanes, p->gen_lanes,
                       precision == MTS_FP32 ? (int)sizeof(float) : (int)sizeof(double),
                       p->tiles[d][0], p->tiles[d][1]);
    }
    // exactly what instance_from_plan allocates
    size_t b = p->field_elems * ((precision == MTS_FP32 ? sizeof(float) : sizeof(double)) +
                                 (method == MTS_ADI ? (precision == MTS_FP32 ? sizeof(float) : sizeof(double)) : 0) +
                                 sizeof(int8_t));
    for (int d = 0; d < 3; d++)
      b += t.tab[d].size() * sizeof(double) + t.blk_ptr[d].size() * sizeof(int) +
           t.blk_hdr[d].size() * sizeof(mts::BlockHdr) + t.blk_val[d].size() * sizeof(double) +
           t.cr_val[d].size() * sizeof(double) + (t.n_seq[d] > 0 ? t.seq[d].size() : 0) +
           t.cls0[d].size();
    for (auto &r : p->recs) b += r.second.size() * sizeof(double);
    for (int d = 0; d < 3; d++)
      b += p->reg[d][0].size() + p->reg[d][1].size() +
           (p->tiles[d][0].size() + p->tiles[d][1].size()) * sizeof(int);
    b += t.spec.size() * sizeof(int);
    mts::DevGeom dg;
    for (int a = 0; a < 3; a++) dg.n[a] = g.n[a];
    b += (size_t)mts::norms_blocks(dg) * (2 * sizeof(double) + sizeof(long long));
    p->device_bytes = b;
    out = std::move(p);
  } catch (const std::bad_alloc &) {
    return fail(MTS_E_NOMEM, "out of host memory planning the instance");
  }
  return MTS_OK;
}

int mts::instance_from_plan(mts_ctx *ctx, mts::PlanPtr plan, mts_instance **out) {
  if (!ctx || !plan || !out) return fail(MTS_E_ARG, "null argument");
  *out = nullptr;
  mts::InstancePlan &p = *plan;
  const mts::Grid &g = p.grid;
  for (int d = 0; d < 3; d++) {
    mts::LinePart q;
    mts::make_line_part(g.n[d], q, p.min_lanes);
    const size_t need =
        ((size_t)mts::tile_lines(q.L) * (q.lds_len * 2 + mts::kWtLen) + 2 * (size_t)g.n[2]) *
            sizeof(double) +
        16 + 2 * (size_t)mts::reg_side_len(q.L, q.mmax) * sizeof(double);  // (the larger kernel)
    if (need > ctx->prop.sharedMemPerBlock)
      return fail(MTS_E_ARG, "line too long for the workgroup LDS of this device");
  }
  MTS_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<mts_instance> in(new mts_instance());
  in->ctx = ctx;
  in->method = p.method;
  in->precision = p.precision;
  in->dt = p.dt;
  in->theta_dt = p.theta_dt;
  in->grid = g;
  in->min_lanes = p.min_lanes;
  in->gen_lanes = p.gen_lanes;
  if (const char *ev = std::getenv("MTS_DOUGLAS_FUSED")) in->douglas_fused = std::atoi(ev) != 0;
  in->px = p.px;
  in->field_elems = p.field_elems;
  MTS_HIP(hipStreamCreateWithFlags(&in->sh.s, hipStreamNonBlocking));
  in->stream = in->sh.s;
  // MTS_SIDE_STREAM=1: the general and the regular tiles of a solve stage run
  // side by side on two streams (default: one after the other on the
  // instance's stream; the fork and join cost more than the overlap gains
  // outside graphs)
  const char *ss = std::getenv("MTS_SIDE_STREAM");
  if (ss && std::atoi(ss) == 1) {
    MTS_HIP(hipStreamCreateWithFlags(&in->side_sh.s, hipStreamNonBlocking));
    in->side.s = in->side_sh.s;
    MTS_HIP(hipEventCreateWithFlags(&in->side.fork, hipEventDisableTiming));
    MTS_HIP(hipEventCreateWithFlags(&in->side.join, hipEventDisableTiming));
  }

  const mts::DeviceTableHost &t = p.t;
  in->phys = t.phys;

  // Every fill and copy below goes on the instance's own stream, which is
  // non-blocking: work on the legacy null stream (hipMemset, and the staged
  // DMA of a pageable hipMemcpy, which may still be in flight when hipMemcpy
  // returns) is not ordered before the kernels launched on it.  The plan's
  // host staging stays alive until the stream is synchronized at the end.
  const hipStream_t st = in->stream;
  if (in->f32()) {
    MTS_HIP(in->uf.alloc(in->field_elems, st));
    MTS_HIP(hipMemsetAsync(in->uf.p, 0, in->field_elems * sizeof(float), st));
  } else {
    MTS_HIP(in->u.alloc(in->field_elems, st));
    MTS_HIP(hipMemsetAsync(in->u.p, 0, in->field_elems * sizeof(double), st));
  }
  if (p.method == MTS_ADI && in->f32()) {
    MTS_HIP(in->rf.alloc(in->field_elems, st));
    MTS_HIP(hipMemsetAsync(in->rf.p, 0, in->field_elems * sizeof(float), st));
  } else if (p.method == MTS_ADI) {
    MTS_HIP(in->r.alloc(in->field_elems, st));
    MTS_HIP(hipMemsetAsync(in->r.p, 0, in->field_elems * sizeof(double), st));
  }
  MTS_HIP(in->ind.upload(p.ind.data(), p.ind.size(), st));
  for (int a = 0; a < 3; a++) MTS_HIP(in->tab[a].upload(t.tab[a].data(), t.tab[a].size(), st));
  for (int d = 0; d < 3; d++) {
    MTS_HIP(in->blk_ptr[d].upload(t.blk_ptr[d].data(), t.blk_ptr[d].size(), st));
    MTS_HIP(in->blk_hdr[d].upload(t.blk_hdr[d].data(), t.blk_hdr[d].size(), st));
    MTS_HIP(in->blk_val[d].upload(t.blk_val[d].data(), t.blk_val[d].size(), st));
    MTS_HIP(in->cr_val[d].upload(t.cr_val[d].data(), t.cr_val[d].size(), st));
    if (t.n_seq[d] > 0) MTS_HIP(in->seq[d].upload(t.seq[d].data(), t.seq[d].size(), st));
    MTS_HIP(in->cls0[d].upload(t.cls0[d].data(), t.cls0[d].size(), st));
  }
  for (int d = 0; d < 3; d++) {
    in->reg_side[d] = p.reg_side[d];
    for (int f = 0; f < 2; f++) {
      MTS_HIP(in->regt[d][f].upload(p.reg[d][f].data(), p.reg[d][f].size(), st));
      MTS_HIP(in->tiles[d][f].upload(p.tiles[d][f].data(), p.tiles[d][f].size(), st));
    }
  }
  for (auto &r : p.recs) {
    std::unique_ptr<DevBuf<double>> b(new DevBuf<double>());
    MTS_HIP(b->upload(r.second.data(), r.second.size(), st));
    in->recs[r.first] = std::move(b);
  }
  in->nspec = (int)(t.spec.size() / mts::kSpecLen);
  MTS_HIP(in->spec.upload(t.spec.data(), t.spec.size(), st));
  mts::DevGeom &dg = in->geom;
  for (int a = 0; a < 3; a++) {
    dg.n[a] = g.n[a];
    dg.tab[a] = in->tab[a].p;
  }
  dg.px = in->px;
  dg.ind = in->ind.p;
  for (int d = 0; d < 3; d++) {
    in->mib[d].blk_ptr = in->blk_ptr[d].p;
    in->mib[d].hdr = in->blk_hdr[d].p;
    in->mib[d].val = in->blk_val[d].p;
    in->mib[d].cr = in->cr_val[d].p;
    in->mib[d].seq = in->seq[d].p;
    in->mib[d].cls0 = in->cls0[d].p;
  }
  int nb = mts::norms_blocks(dg);
  MTS_HIP(in->part_max.alloc(nb, st));
  MTS_HIP(in->part_sq.alloc(nb, st));
  MTS_HIP(in->part_idx.alloc(nb, st));
  MTS_HIP(hipStreamSynchronize(st));
  *out = in.release();
  return MTS_OK;
}

extern "C" {

const char *mts_last_error(void) { return g_err.c_str(); }
const char *mts_version(void) { return "mts-amd 0.1 (gfx950)"; }

int mts_setup_reference_stops(const mts_setup *s, mts_reference_stops *out) {
  if (!s || !out) return fail(MTS_E_ARG, "null argument");
  mts::RefStops r;
  mts::reference_stops(s->res, r);
  std::memset(out, 0, sizeof(*out));
  for (int k = 0; k < 4; k++) out->count[k] = r.count[k];
  out->first_kind = r.first_kind;
  out->first_dir = r.first_dir;
  out->first_index = r.first_index;
  std::snprintf(out->message, sizeof out->message, "%s", r.message.c_str());
  return MTS_OK;
}

int mts_setup_create(const mts_grid_desc *grid, const mts_surface_desc *surface,
                     const mts_physics_desc *physics, int mib, int accuracy, mts_setup **out) {
  if (!grid || !surface || !physics || !out) return fail(MTS_E_ARG, "null argument");
  *out = nullptr;
  for (int a = 0; a < 3; a++) {
    if (grid->n[a] < 5) return fail(MTS_E_ARG, "each axis needs at least 5 nodes");
    if (!(grid->hi[a] > grid->lo[a])) return fail(MTS_E_ARG, "empty domain");
  }
  std::string err;
  std::unique_ptr<mts::Surface> surf = make_surface(surface, err);
  if (!surf) return fail(MTS_E_ARG, err);
  std::unique_ptr<mts_setup> s(new (std::nothrow) mts_setup());
  if (!s) return fail(MTS_E_NOMEM, "out of host memory");
  if (!make_beta(physics, s->beta, err)) return fail(MTS_E_ARG, err);
  mts::Grid g = mts::make_grid(grid->lo, grid->hi, grid->n);
  mts::BuildOptions opt;
  opt.mib = mib;
  opt.accuracy = accuracy;
  try {
    err = mts::build_setup(g, *surf, s->beta, opt, s->res);
  } catch (const std::bad_alloc &) {
    return fail(MTS_E_NOMEM, "out of host memory in setup");
  }
  if (!err.empty()) return fail(MTS_E_GEOMETRY, err);
  *out = s.release();
  return MTS_OK;
}

void mts_setup_destroy(mts_setup *s) { delete s; }

int mts_setup_diagnostics(const mts_setup *s, double t, double fp[2], double jump[2]) {
  if (!s || !fp || !jump) return fail(MTS_E_ARG, "null argument");
  mts::interface_diagnostics(s->res, s->beta, t, fp, jump);
  return MTS_OK;
}

int mts_setup_crossing_errors(const mts_setup *s, int dir, double t, double *errl, double *errr,
                              double *jerr, size_t count) {
  if (!s || !errl || !errr || !jerr) return fail(MTS_E_ARG, "null argument");
  if (dir < 0 || dir > 2) return fail(MTS_E_ARG, "dir must be 0, 1 or 2");
  if (count < s->res.cross[dir].size()) return fail(MTS_E_ARG, "output smaller than the crossing count");
  std::vector<double> l, r, j;
  mts::crossing_errors(s->res, s->beta, dir, t, l, r, j);
  std::copy(l.begin(), l.end(), errl);
  std::copy(r.begin(), r.end(), errr);
  std::copy(j.begin(), j.end(), jerr);
  return MTS_OK;
}

int mts_mesh_classify(const mts_grid_desc *grid, const mts_surface_desc *surface, int8_t *out,
                      size_t count) {
  if (!grid || !surface || !out) return fail(MTS_E_ARG, "null argument");
  for (int a = 0; a < 3; a++) {
    if (grid->n[a] < 2) return fail(MTS_E_ARG, "each axis needs at least 2 nodes");
    if (!(grid->hi[a] > grid->lo[a])) return fail(MTS_E_ARG, "empty domain");
  }
  if (count < (size_t)grid->n[0] * grid->n[1] * grid->n[2])
    return fail(MTS_E_ARG, "output smaller than the node count");
  std::string err;
  std::unique_ptr<mts::Surface> surf = make_surface(surface, err);
  if (!surf) return fail(MTS_E_ARG, err);
  const mts::Grid g = mts::make_grid(grid->lo, grid->hi, grid->n);
  std::vector<int8_t> ind;
  try {
    err = mts::classify_nodes(g, *surf, ind);
  } catch (const std::bad_alloc &) {
    return fail(MTS_E_NOMEM, "out of host memory");
  }
  if (!err.empty()) return fail(MTS_E_GEOMETRY, err);
  std::memcpy(out, ind.data(), ind.size());
  return MTS_OK;
}

int mts_lu_create(int n, const double *a, mts_lu **out) {
  if (n < 1 || !a || !out) return fail(MTS_E_ARG, "null argument or n < 1");
  *out = nullptr;
  std::unique_ptr<mts_lu> l(new (std::nothrow) mts_lu(n, a));
  if (!l) return fail(MTS_E_NOMEM, "out of host memory");
  // lu.cc:32-35: a row of zeros has no scale
  if (l->lu.singular_row()) return fail(MTS_E_ARG, "Singular matrix in LUdcmp");
  *out = l.release();
  return MTS_OK;
}

void mts_lu_destroy(mts_lu *l) { delete l; }

int mts_lu_solve(const mts_lu *l, const double *b, double *x) {
  if (!l || !b || !x) return fail(MTS_E_ARG, "null argument");
  l->lu.solve(b, x);
  return MTS_OK;
}

int mts_lu_solve_matrix(const mts_lu *l, int m, const double *b, double *x) {
  if (!l || !b || !x || m < 1) return fail(MTS_E_ARG, "null argument or m < 1");
  const int n = l->lu.n();
  std::vector<double> col(n);
  for (int j = 0; j < m; j++) {  // column by column (lu.cc:124-143)
    for (int i = 0; i < n; i++) col[i] = b[(size_t)i * m + j];
    l->lu.solve(col.data(), col.data());
    for (int i = 0; i < n; i++) x[(size_t)i * m + j] = col[i];
  }
  return MTS_OK;
}

int mts_lu_inverse(const mts_lu *l, double *ainv) {
  if (!l || !ainv) return fail(MTS_E_ARG, "null argument");
  l->lu.inverse(ainv);
  return MTS_OK;
}

int mts_lu_det(const mts_lu *l, double *det) {
  if (!l || !det) return fail(MTS_E_ARG, "null argument");
  *det = l->lu.det();
  return MTS_OK;
}

int mts_lu_mprove(const mts_lu *l, const double *b, double *x) {
  if (!l || !b || !x) return fail(MTS_E_ARG, "null argument");
  const int n = l->lu.n();
  // residual A x - b accumulated in extended precision (lu.cc:179-196)
  std::vector<double> r(n);
  for (int i = 0; i < n; i++) {
    long double sdp = -(long double)b[i];
    for (int j = 0; j < n; j++) sdp += (long double)l->a[(size_t)i * n + j] * (long double)x[j];
    r[i] = (double)sdp;
  }
  l->lu.solve(r.data(), r.data());
  for (int i = 0; i < n; i++) x[i] -= r[i];
  return MTS_OK;
}

int mts_setup_get_info(const mts_setup *s, mts_setup_info *info) {
  if (!s || !info) return fail(MTS_E_ARG, "null argument");
  for (int a = 0; a < 3; a++) {
    info->n[a] = s->res.grid.n[a];
    info->n_cross[a] = (long long)s->res.cross[a].size();
    info->corner_pairs[a] = s->res.corner_pairs[a];
  }
  info->singular_warnings = s->res.singular_warnings;
  info->mib = s->res.mib;
  return MTS_OK;
}

int mts_setup_indicator(const mts_setup *s, int8_t *out, size_t count) {
  if (!s || !out) return fail(MTS_E_ARG, "null argument");
  if (count < s->res.ind.size()) return fail(MTS_E_ARG, "indicator buffer too small");
  std::memcpy(out, s->res.ind.data(), s->res.ind.size());
  return MTS_OK;
}

int mts_setup_coords(const mts_setup *s, int axis, double *out, size_t count) {
  if (!s || !out || axis < 0 || axis > 2) return fail(MTS_E_ARG, "bad argument");
  const std::vector<double> &c = s->res.grid.c[axis];
  if (count < c.size()) return fail(MTS_E_ARG, "coordinate buffer too small");
  std::memcpy(out, c.data(), c.size() * sizeof(double));
  return MTS_OK;
}

int mts_setup_crossings(const mts_setup *s, int dir, mts_crossing *out, size_t count) {
  if (!s || !out || dir < 0 || dir > 2) return fail(MTS_E_ARG, "bad argument");
  const std::vector<mts::Crossing> &cs = s->res.cross[dir];
  if (count < cs.size()) return fail(MTS_E_ARG, "crossing buffer too small");
  for (size_t i = 0; i < cs.size(); i++) {
    const mts::Crossing &c = cs[i];
    mts_crossing &o = out[i];
    std::memset(&o, 0, sizeof(o));
    o.dir = c.dir;
    o.j = c.j;
    o.k = c.k;
    o.id = c.id;
    o.left_loc = c.left_loc;
    o.nw = c.nw;
    o.gamma = c.gamma;
    for (int a = 0; a < 3; a++) {
      o.coord[a] = c.coord[a];
      o.normal[a] = c.normal[a];
      for (int b = 0; b < 3; b++) o.p[a * 3 + b] = c.p[a][b];
    }
    o.pinv_d0 = c.pinv_d0;
    o.beta_in = c.beta_in;
    o.beta_out = c.beta_out;
    for (int w = 0; w < 9; w++) {
      o.wl[w] = c.wl[w];
      o.wr[w] = c.wr[w];
    }
  }
  return MTS_OK;
}

int mts_ctx_create(int device, mts_ctx **out) {
  if (!out) return fail(MTS_E_ARG, "null argument");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0)
    return fail(MTS_E_DEVICE, "no HIP device available (the per-step path runs on the GPU only)");
  if (device < 0 || device >= count) return fail(MTS_E_ARG, "device index out of range");
  std::unique_ptr<mts_ctx> c(new mts_ctx());
  c->device = device;
  MTS_HIP(hipSetDevice(device));
  MTS_HIP(hipGetDeviceProperties(&c->prop, device));
  MTS_HIP(mts::norms_scratch_ptr(device, &c->norms_scratch));
  *out = c.release();
  return MTS_OK;
}

void mts_ctx_destroy(mts_ctx *ctx) { delete ctx; }

int mts_ctx_device_info(mts_ctx *ctx, char *name, size_t name_len, size_t *hbm_bytes) {
  if (!ctx) return fail(MTS_E_ARG, "null context");
  if (name && name_len) std::snprintf(name, name_len, "%s (%s)", ctx->prop.name, ctx->prop.gcnArchName);
  if (hbm_bytes) *hbm_bytes = ctx->prop.totalGlobalMem;
  return MTS_OK;
}

int mts_instance_create(mts_ctx *ctx, const mts_setup *setup, const mts_physics_desc *physics,
                        int method, double dt, mts_instance **out) {
  if (!ctx || !out) return fail(MTS_E_ARG, "null argument");
  *out = nullptr;
  return mts_instance_create_ex(ctx, setup, physics, method, dt, MTS_FP64, out);
}

int mts_instance_create_ex(mts_ctx *ctx, const mts_setup *setup, const mts_physics_desc *physics,
                           int method, double dt, int precision, mts_instance **out) {
  if (!ctx || !out) return fail(MTS_E_ARG, "null argument");
  *out = nullptr;
  mts::PlanPtr plan;
  mts::Lanes lanes;
  const bool tuned =
      setup && mts::ctx_tuned_lanes(ctx, setup->res.grid.n, method, precision, lanes);
  int rc = mts::plan_instance(setup, physics, method, dt, precision, plan, tuned ? &lanes : nullptr);
  if (rc) return rc;
  return mts::instance_from_plan(ctx, std::move(plan), out);
}

int mts_instance_lanes(const mts_instance *in, int *min_lanes, int *gen_lanes) {
  if (!in || !min_lanes || !gen_lanes) return fail(MTS_E_ARG, "null argument");
  *min_lanes = in->min_lanes;
  *gen_lanes = in->gen_lanes;
  return MTS_OK;
}

int mts_tune_lanes(mts_ctx *ctx, const mts_setup *setup, const mts_physics_desc *physics,
                   int method, double dt, int precision, int steps, int *min_lanes,
                   int *gen_lanes, double *ms_per_step) {
  if (!ctx || !setup || !physics || !min_lanes || !gen_lanes)
    return fail(MTS_E_ARG, "null argument");
  if (steps < 1 || steps > 1000) return fail(MTS_E_ARG, "steps must be within 1..1000");
  static const int kCand[6][2] = {{16, 16}, {16, 32}, {16, 64}, {32, 32}, {32, 64}, {64, 64}};
  const int *n = setup->res.grid.n;
  mts::Lanes best;
  double best_ms = 0.0;
  int last_rc = MTS_OK;
  for (const auto &c : kCand) {
    // splits that do not exist for these lines (longer than 16 or 32 lanes
    // can hold) are skipped
    bool ok = true;
    for (int d = 0; d < 3 && ok; d++) {
      mts::LinePart q, qg;
      ok = mts::make_line_part(n[d], q, c[0]) && q.L == c[0] &&
           mts::make_line_part(n[d], qg, c[1]);
    }
    if (!ok) continue;
    mts::Lanes l;
    l.min_lanes = c[0];
    l.gen_lanes = c[1];
    mts::PlanPtr plan;
    int rc = mts::plan_instance(setup, physics, method, dt, precision, plan, &l);
    mts_instance *in = nullptr;
    if (!rc) rc = mts::instance_from_plan(ctx, std::move(plan), &in);
    if (rc) {
      last_rc = rc;
      if (rc == MTS_E_DISABLED || rc == MTS_E_ARG) return rc;  // (the same for every split)
      continue;
    }
    // from the initial field: two untimed steps, then `steps` timed ones
    double t = 0.0;
    rc = mts_initialize(in, 0.0);
    for (int k = 0; k < 2 && !rc; k++) rc = mts_step(in, t += dt);
    if (!rc) rc = mts_sync(in);
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < steps && !rc; k++) rc = mts_step(in, t += dt);
    if (!rc) rc = mts_sync(in);
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
        steps;
    mts_instance_destroy(in);
    if (rc) return rc;
    if (!best.min_lanes || ms < best_ms) {
      best = l;
      best_ms = ms;
    }
  }
  if (!best.min_lanes) return last_rc ? last_rc : fail(MTS_E_ARG, "no line split fits this grid");
  {
    std::lock_guard<std::mutex> lk(ctx->tune_mu);
    ctx->tuned[{n[0], n[1], n[2], method, precision}] = best;
  }
  *min_lanes = best.min_lanes;
  *gen_lanes = best.gen_lanes;
  if (ms_per_step) *ms_per_step = best_ms;
  return MTS_OK;
}

int mts_instance_precision(const mts_instance *in, int *precision) {
  if (!in || !precision) return fail(MTS_E_ARG, "null argument");
  *precision = in->precision;
  return MTS_OK;
}

void mts_instance_destroy(mts_instance *inst) {
  if (!inst) return;
  (void)hipSetDevice(inst->ctx->device);
  (void)hipStreamSynchronize(inst->stream);
  delete inst;
}

static void pack_field(const mts_instance *in, const double *host, int layout,
                       std::vector<double> &buf) {
  const mts::Grid &g = in->grid;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  buf.assign(in->field_elems, 0.0);
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++) {
      double *row = &buf[((size_t)iz * ny + iy) * in->px];
      if (layout == MTS_LAYOUT_XFAST) {
        std::memcpy(row, host + ((size_t)iz * ny + iy) * nx, nx * sizeof(double));
      } else {
        for (int ix = 0; ix < nx; ix++) row[ix] = host[((size_t)ix * ny + iy) * nz + iz];
      }
    }
}

int mts_field_upload(mts_instance *in, const double *host, int layout) {
  if (!in || !host) return fail(MTS_E_ARG, "null argument");
  if (layout != MTS_LAYOUT_XFAST && layout != MTS_LAYOUT_REF) return fail(MTS_E_ARG, "bad layout");
  MTS_HIP(hipSetDevice(in->ctx->device));
  std::vector<double> buf;
  pack_field(in, host, layout, buf);
  if (in->f32()) {
    std::vector<float> bf(buf.begin(), buf.end());
    MTS_HIP(hipMemcpyAsync(in->uf.p, bf.data(), bf.size() * sizeof(float), hipMemcpyHostToDevice,
                           in->stream));
    MTS_HIP(hipStreamSynchronize(in->stream));
    return MTS_OK;
  }
  MTS_HIP(hipMemcpyAsync(in->u.p, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice,
                         in->stream));
  MTS_HIP(hipStreamSynchronize(in->stream));
  return MTS_OK;
}

int mts_field_download(mts_instance *in, double *host, int layout) {
  if (!in || !host) return fail(MTS_E_ARG, "null argument");
  if (layout != MTS_LAYOUT_XFAST && layout != MTS_LAYOUT_REF) return fail(MTS_E_ARG, "bad layout");
  MTS_HIP(hipSetDevice(in->ctx->device));
  std::vector<double> buf(in->field_elems);
  if (in->f32()) {
    std::vector<float> bf(in->field_elems);
    MTS_HIP(hipMemcpyAsync(bf.data(), in->uf.p, bf.size() * sizeof(float), hipMemcpyDeviceToHost,
                           in->stream));
    MTS_HIP(hipStreamSynchronize(in->stream));
    std::copy(bf.begin(), bf.end(), buf.begin());
  } else {
    MTS_HIP(hipMemcpyAsync(buf.data(), in->u.p, buf.size() * sizeof(double),
                           hipMemcpyDeviceToHost, in->stream));
    MTS_HIP(hipStreamSynchronize(in->stream));
  }
  const mts::Grid &g = in->grid;
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  for (int iz = 0; iz < nz; iz++)
    for (int iy = 0; iy < ny; iy++) {
      const double *row = &buf[((size_t)iz * ny + iy) * in->px];
      if (layout == MTS_LAYOUT_XFAST) {
        std::memcpy(host + ((size_t)iz * ny + iy) * nx, row, nx * sizeof(double));
      } else {
        for (int ix = 0; ix < nx; ix++) host[((size_t)ix * ny + iy) * nz + iz] = row[ix];
      }
    }
  return MTS_OK;
}

int mts_initialize(mts_instance *in, double t0) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  MTS_HIP(hipSetDevice(in->ctx->device));
  MTS_HIP(mts::launch_init(in->geom, in->field(), in->f32(), std::exp(-t0), in->stream));
  return MTS_OK;
}

static mts::LineArgs base_args(mts_instance *in) {
  mts::LineArgs a;
  std::memset(&a, 0, sizeof(a));
  a.min_lanes = in->min_lanes;
  a.g = in->geom;
  a.P = in->phys;
  a.theta_dt = in->theta_dt;
  a.dt = in->dt;
  return a;
}

// One line stage with the stage records of its direction and kind.
static hipError_t line_launch(mts_instance *in, int dir, int flags, const mts::LineArgs &args,
                              hipStream_t st) {
  mts::LineArgs a = args;
  a.mib.rec = in->rec_for(dir, flags);
  if (!a.mib.rec && in->mib[dir].hdr) return hipErrorInvalidValue;
  const bool solve = (flags & mts::kFlagSolve) != 0;
  a.reg[0] = solve ? in->regt[dir][0].p : nullptr;
  a.reg[1] = solve ? in->regt[dir][1].p : nullptr;
  a.reg_side = in->reg_side[dir];
  a.reg_sides = in->phys.beta_kind == 0 ? 2 : 1;
  // the stage's tiles split into the regular and the other ones (an empty
  // list has no buffer and a count of 0)
  a.split = solve ? 1 : 0;
  a.gen_lanes = in->gen_lanes;
  for (int c = 0; c < 2; c++) {
    a.tiles[c] = solve ? in->tiles[dir][c].p : nullptr;
    a.ntiles[c] = solve ? (int)in->tiles[dir][c].n : 0;
  }
  if (in->rec) {
    mts_instance::StepOp op;
    op.dir = dir;
    op.flags = flags;
    op.a = a;
    if (solve) {
      op.line = 1;
    } else {
      const bool f32 = in->f32();
      op.fn = [dir, flags, f32, a](hipStream_t s) { return mts::launch_line(dir, flags, f32, a, s); };
    }
    in->rec->push_back(std::move(op));
    return hipSuccess;
  }
  return mts::launch_line(dir, flags, in->f32(), a, st, &in->side);
}

// The boundary faces at time factor e (with the source term's cf): run, or
// recorded for a launch over several instances.
static hipError_t faces_op(mts_instance *in, double e, double cf) {
  if (in->rec) {
    mts_instance::StepOp op;
    op.line = 2;
    op.fa.g = in->geom;
    op.fa.P = in->phys;
    op.fa.out = in->field();
    op.fa.e = e;
    op.fa.cf = cf;
    in->rec->push_back(std::move(op));
    return hipSuccess;
  }
  return timed(in, "faces", [&] {
    return mts::launch_faces(in->geom, in->phys, in->field(), in->f32(), e, cf, in->stream);
  });
}
// Douglas stage 1: run, or recorded.
static hipError_t rhs1_op(mts_instance *in, const char *name, const mts::Rhs1Args &r1,
                          const mts::SpecArgs &sp, bool pairs_too) {
  if (in->rec) {
    mts_instance::StepOp op;
    op.line = 3;
    op.r1 = r1;
    op.sp = sp;
    op.pairs_too = pairs_too;
    in->rec->push_back(std::move(op));
    return hipSuccess;
  }
  return timed(in, name,
               [&] { return mts::launch_rhs1(r1, sp, pairs_too, in->f32(), in->stream); });
}


int mts_step(mts_instance *in, double t_next) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  MTS_HIP(hipSetDevice(in->ctx->device));
  hipStream_t st = in->stream;
  const double dt = in->dt;
  const double e1 = std::exp(-t_next);
  using namespace mts;
  if (in->method == MTS_ADI) {
    Rhs1Args r1;
    r1.g = in->geom;
    r1.P = in->phys;
    r1.u = in->field();
    r1.R = in->stage();
    r1.dt = dt;
    r1.e = e1;
    SpecArgs sp;
    sp.g = in->geom;
    sp.P = in->phys;
    for (int d = 0; d < 3; d++) sp.mib[d] = in->mib[d];
    sp.spec = in->spec.p;
    sp.nspec = in->nspec;
    sp.u = in->field();
    sp.R = in->stage();
    sp.dt = dt;
    sp.e = e1;
    // Fused (the default): the x sweep computes the stage-1 right-hand side
    // of u^n as it stages its tiles (kFlagPre1), so only the MIB-affected
    // nodes are computed beforehand (into r, where the x sweep reads them) and
    // R is never written out and read back.  MTS_DOUGLAS_FUSED=0, and the
    // FP32 form: the whole right-hand side into r first, then the x sweep
    // over r.
    const bool fused = in->douglas_fused && !in->f32();
    MTS_HIP(rhs1_op(in, fused ? "adi_spec" : "adi_rhs1", r1, sp, !fused));
    LineArgs a = base_args(in);
    a.e_out = e1;
    a.e_jump = e1;
    a.mib = in->mib[0];
    a.A = fused ? in->field() : in->stage();
    a.S = fused ? in->stage() : nullptr;
    a.out = in->stage();
    a.in_analytic = fused ? 1 : 0;  // the x sweep's end rows: boundary values at t^{n+1}
    a.e_in = e1;
    MTS_HIP(timed(in, "adi_x", [&] {
      return line_launch(in, 0, kRhsCopy | kFlagSolve | (fused ? kFlagPre1 : 0), a, st);
    }));
    a.in_analytic = 0;
    a.A = in->stage();
    a.mib = in->mib[1];
    a.S = in->field();
    MTS_HIP(timed(in, "adi_y", [&] { return line_launch(in, 1, kRhsDouglas | kFlagSolve, a, st); }));
    a.mib = in->mib[2];
    a.out = in->field();
    MTS_HIP(timed(in, "adi_z", [&] { return line_launch(in, 2, kRhsDouglas | kFlagSolve, a, st); }));
    MTS_HIP(faces_op(in, e1, 0.0));
  } else if (in->method == MTS_LOD_IE || in->method == MTS_LOD_CN) {
    const int rhs = in->method == MTS_LOD_CN ? kRhsCN : kRhsCopy;
    LineArgs a = base_args(in);
    a.e_out = e1;
    a.e_jump = e1;
    a.A = in->field();
    a.out = in->field();
    a.mib = in->mib[0];
    MTS_HIP(timed(in, "lod_x", [&] { return line_launch(in, 0, rhs | kFlagSolve | kFlagLhsJump, a, st); }));
    a.in_analytic = 1;
    a.e_in = e1;
    a.mib = in->mib[1];
    MTS_HIP(timed(in, "lod_y", [&] { return line_launch(in, 1, rhs | kFlagSolve | kFlagLhsJump, a, st); }));
    a.mib = in->mib[2];
    a.e_f1 = e1;
    a.c_f1 = dt;
    MTS_HIP(timed(in, "lod_z", [&] {
      return line_launch(in, 2, rhs | kFlagSolve | kFlagLhsJump | kFlagEpiF, a, st);
    }));
    MTS_HIP(faces_op(in, e1, dt));
  } else {
    const double en = std::exp(-(t_next - dt));
    const double eh = std::exp(-(t_next - 0.5 * dt));
    LineArgs a = base_args(in);
    a.e_jump = e1;
    a.A = in->field();
    a.out = in->field();
    // explicit (I + dt/2 L_x), (I + dt/2 L_y) with boundary values at t^{n+1/2}
    a.e_out = eh;
    a.mib = in->mib[0];
    MTS_HIP(timed(in, "ts_ex", [&] { return line_launch(in, 0, kRhsCN, a, st); }));
    a.in_analytic = 1;
    a.e_in = eh;
    a.mib = in->mib[1];
    MTS_HIP(timed(in, "ts_ey", [&] { return line_launch(in, 1, kRhsCN, a, st); }));
    // explicit z, sources at t^n and t^{n+1}, implicit z
    a.mib = in->mib[2];
    a.e_out = e1;
    a.e_f1 = en;
    a.c_f1 = 0.5 * dt;
    a.e_f2 = e1;
    a.c_f2 = 0.5 * dt;
    MTS_HIP(timed(in, "ts_z", [&] { return line_launch(in, 2, kRhsTSZ | kFlagSolve | kFlagLhsJump, a, st); }));
    // implicit y, x
    a.e_in = e1;
    a.mib = in->mib[1];
    MTS_HIP(timed(in, "ts_iy", [&] { return line_launch(in, 1, kRhsCopy | kFlagSolve | kFlagLhsJump, a, st); }));
    a.mib = in->mib[0];
    MTS_HIP(timed(in, "ts_ix", [&] { return line_launch(in, 0, kRhsCopy | kFlagSolve | kFlagLhsJump, a, st); }));
    MTS_HIP(faces_op(in, e1, 0.0));
  }
  return MTS_OK;
}

int mts_batch_run_multi(mts_instance *const *insts, int count, const double *t0,
                        const double *t_finish, int *steps, double *t_end) {
  if (!insts || count < 1 || !t0 || !t_finish) return fail(MTS_E_ARG, "null argument");
  for (int i = 0; i < count; i++) {
    if (!insts[i]) return fail(MTS_E_ARG, "null instance");
    if (insts[i]->ctx != insts[0]->ctx) return fail(MTS_E_ARG, "instances of different contexts");
  }
  MTS_HIP(hipSetDevice(insts[0]->ctx->device));
  // every instance's earlier work is done (their buffers are then usable on
  // the batch stream)
  for (int i = 0; i < count; i++) MTS_HIP(hipStreamSynchronize(insts[i]->stream));
  const hipStream_t s = insts[0]->stream;
  std::vector<int> loops(count);
  std::vector<double> t(count);
  for (int i = 0; i < count; i++) {
    loops[i] = (int)(t_finish[i] / insts[i]->dt);
    t[i] = t0[i];
  }
  // groups of one method and precision, in order of first appearance
  std::vector<std::vector<int>> groups;
  for (int i = 0; i < count; i++) {
    size_t g = 0;
    while (g < groups.size() && !(insts[groups[g][0]]->method == insts[i]->method &&
                                  insts[groups[g][0]]->precision == insts[i]->precision))
      g++;
    if (g == groups.size()) groups.emplace_back();
    groups[g].push_back(i);
  }
  mts::MultiBuf buf;
  std::vector<std::vector<mts_instance::StepOp>> ops(count);
  std::vector<mts::LineArgs> args;
  int rc = MTS_OK;
  for (const auto &grp : groups) {
    int maxl = 0;
    for (int i : grp) maxl = std::max(maxl, loops[i]);
    for (int k = 0; k < maxl && rc == MTS_OK; k++) {
      std::vector<int> act;
      for (int i : grp)
        if (k < loops[i]) act.push_back(i);
      for (int i : act) {
        mts_instance *in = insts[i];
        ops[i].clear();
        const bool prof = in->profile;
        in->profile = false;
        in->rec = &ops[i];
        t[i] += in->dt;
        rc = mts_step(in, t[i]);
        in->rec = nullptr;
        in->profile = prof;
        if (rc) break;
      }
      if (rc) break;
      const size_t nops = ops[act[0]].size();
      for (int i : act)
        if (ops[i].size() != nops) {
          rc = fail(MTS_E_ARG, "instances of one method recorded different steps");
          break;
        }
      for (size_t o = 0; o < nops && rc == MTS_OK; o++) {
        const mts_instance::StepOp &first = ops[act[0]][o];
        bool same_kind = true;
        for (int i : act) same_kind &= ops[i][o].line == first.line;
        if (!same_kind) {
          rc = fail(MTS_E_ARG, "instances of one method recorded different steps");
          break;
        }
        if (first.line == 2) {
          std::vector<mts::FaceArgs> fa;
          for (int i : act) fa.push_back(ops[i][o].fa);
          const hipError_t e = mts::launch_faces_multi(fa.data(), (int)fa.size(),
                                                       insts[act[0]]->f32(), buf, s);
          if (e != hipSuccess) rc = fail(MTS_E_DEVICE, hipGetErrorString(e));
        } else if (first.line == 3) {
          std::vector<mts::Rhs1Args> r1;
          std::vector<mts::SpecArgs> sp;
          for (int i : act) {
            if (ops[i][o].pairs_too != first.pairs_too) {
              rc = fail(MTS_E_ARG, "instances of one method recorded different steps");
              break;
            }
            r1.push_back(ops[i][o].r1);
            sp.push_back(ops[i][o].sp);
          }
          if (rc) break;
          const hipError_t e = mts::launch_rhs1_multi(r1.data(), sp.data(), (int)r1.size(),
                                                      first.pairs_too, insts[act[0]]->f32(), buf, s);
          if (e != hipSuccess) rc = fail(MTS_E_DEVICE, hipGetErrorString(e));
        } else if (first.line == 1) {
          args.clear();
          for (int i : act) {
            const mts_instance::StepOp &op = ops[i][o];
            if (!op.line || op.dir != first.dir || op.flags != first.flags) {
              rc = fail(MTS_E_ARG, "instances of one method recorded different steps");
              break;
            }
            args.push_back(op.a);
          }
          if (rc) break;
          const hipError_t e = mts::launch_line_multi(first.dir, first.flags, insts[act[0]]->f32(),
                                                      args.data(), (int)args.size(), buf, s);
          if (e != hipSuccess) rc = fail(MTS_E_DEVICE, hipGetErrorString(e));
        } else {
          for (int i : act) {
            if (ops[i][o].line) {
              rc = fail(MTS_E_ARG, "instances of one method recorded different steps");
              break;
            }
            const hipError_t e = ops[i][o].fn(s);
            if (e != hipSuccess) {
              rc = fail(MTS_E_DEVICE, hipGetErrorString(e));
              break;
            }
          }
        }
      }
    }
    if (rc) break;
  }
  const hipError_t e = hipStreamSynchronize(s);
  if (rc) return rc;
  if (e != hipSuccess) return fail(MTS_E_DEVICE, hipGetErrorString(e));
  for (int i = 0; i < count; i++) {
    if (steps) steps[i] = loops[i];
    if (t_end) t_end[i] = t[i];
  }
  return MTS_OK;
}

int mts_run(mts_instance *in, double t0, double t_finish, int *steps, double *t_end) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  int loop = (int)(t_finish / in->dt);
  double t = t0;
  for (int j = 0; j < loop; j++) {
    t += in->dt;
    int rc = mts_step(in, t);
    if (rc) return rc;
  }
  if (steps) *steps = loop;
  if (t_end) *t_end = t;
  return MTS_OK;
}

int mts_run_graph(mts_instance *in, double t0, double t_finish, int *steps, double *t_end) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  if (in->profile) return fail(MTS_E_STATE, "graph runs need per-stage profiling off");
  MTS_HIP(hipSetDevice(in->ctx->device));
  const int loop = (int)(t_finish / in->dt);
  if (loop <= 0) {
    if (steps) *steps = 0;
    if (t_end) *t_end = t0;
    return MTS_OK;
  }
  if (!in->gexec || in->g_t0 != t0 || in->g_tf != t_finish) {
    if (in->gexec) {
      MTS_HIP(hipStreamSynchronize(in->stream));
      MTS_HIP(hipGraphExecDestroy(in->gexec));
      in->gexec = nullptr;
    }
    // the steps are recorded exactly as mts_run issues them (kernels only)
    MTS_HIP(hipStreamBeginCapture(in->stream, hipStreamCaptureModeThreadLocal));
    double t = t0;
    int rc = MTS_OK;
    for (int j = 0; j < loop && rc == MTS_OK; j++) {
      t += in->dt;
      rc = mts_step(in, t);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ce = hipStreamEndCapture(in->stream, &graph);
    if (rc != MTS_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (ce != hipSuccess) return fail(MTS_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    const hipError_t ie = hipGraphInstantiate(&in->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
      in->gexec = nullptr;
      return fail(MTS_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    in->g_t0 = t0;
    in->g_tf = t_finish;
    in->g_steps = loop;
    in->g_tend = t;
  }
  MTS_HIP(hipGraphLaunch(in->gexec, in->stream));
  if (steps) *steps = in->g_steps;
  if (t_end) *t_end = in->g_tend;
  return MTS_OK;
}

int mts_sync(mts_instance *in) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  MTS_HIP(hipSetDevice(in->ctx->device));
  MTS_HIP(hipStreamSynchronize(in->stream));
  return MTS_OK;
}

int mts_error_norms(mts_instance *in, double t, double *linf, double *l2, int argmax[3]) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  MTS_HIP(hipSetDevice(in->ctx->device));
  int nparts = 0;
  MTS_HIP(mts::launch_norms(in->geom, in->field(), in->f32(), std::exp(-t), in->part_max.p, in->part_idx.p,
                            in->part_sq.p, &nparts, in->stream));
  std::vector<double> pm(nparts), ps(nparts);
  std::vector<long long> pi(nparts);
  MTS_HIP(hipMemcpyAsync(pm.data(), in->part_max.p, nparts * sizeof(double),
                         hipMemcpyDeviceToHost, in->stream));
  MTS_HIP(hipMemcpyAsync(ps.data(), in->part_sq.p, nparts * sizeof(double),
                         hipMemcpyDeviceToHost, in->stream));
  MTS_HIP(hipMemcpyAsync(pi.data(), in->part_idx.p, nparts * sizeof(long long),
                         hipMemcpyDeviceToHost, in->stream));
  MTS_HIP(hipStreamSynchronize(in->stream));
  double mx = -1.0, sq = 0.0;
  long long mi = 0;
  for (int i = 0; i < nparts; i++) {
    if (pm[i] > mx || (pm[i] == mx && pi[i] < mi)) {
      mx = pm[i];
      mi = pi[i];
    }
    sq += ps[i];
  }
  const mts::Grid &g = in->grid;
  const long long ny = g.n[1], nz = g.n[2];
  if (linf) *linf = mx;
  if (l2) *l2 = std::sqrt(sq / (1.0 * g.n[0] * g.n[1] * g.n[2]));
  if (argmax) {
    argmax[0] = (int)(mi / (ny * nz));
    argmax[1] = (int)((mi / nz) % ny);
    argmax[2] = (int)(mi % nz);
  }
  return MTS_OK;
}

// The interface diagnostics of every crossing at time t into in->diag (el
// at [0, total), er at [total, 2 total), je at [2 total, 3 total), directions
// one after the other); off[d] is direction d's first crossing.
static int diag_run(mts_instance *in, double t, int off[4]) {
  off[0] = 0;
  for (int d = 0; d < 3; d++) off[d + 1] = off[d] + (int)(in->cr_val[d].n / mts::kCrLen);
  const int total = off[3];
  if (total == 0) return MTS_OK;
  if (!in->diag.p) {
    MTS_HIP(in->diag.alloc(3 * (size_t)total + 4, in->stream));
    MTS_HIP(hipMemsetAsync(in->diag.p, 0, (3 * (size_t)total + 4) * sizeof(double), in->stream));
  }
  const mts::Grid &g = in->grid;
  for (int d = 0; d < 3; d++) {
    mts::DiagArgs a;
    a.g = in->geom;
    a.mib = in->mib[d];
    a.cr = in->cr_val[d].p;
    a.nj = d == 0 ? g.n[1] : g.n[0];
    a.nk = d == 2 ? g.n[1] : g.n[2];
    a.e = std::exp(-t);
    a.el = in->diag.p + off[d];
    a.er = in->diag.p + total + off[d];
    a.je = in->diag.p + 2 * (size_t)total + off[d];
    if (off[d + 1] > off[d]) MTS_HIP(mts::launch_diagnostics(d, a, in->stream));
  }
  return MTS_OK;
}

int mts_instance_diagnostics(mts_instance *in, double t, double fp[2], double jump[2]) {
  if (!in || !fp || !jump) return fail(MTS_E_ARG, "null argument");
  MTS_HIP(hipSetDevice(in->ctx->device));
  int off[4];
  int rc = diag_run(in, t, off);
  if (rc) return rc;
  const int total = off[3];
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  if (total > 0) {
    double *base = in->diag.p;
    MTS_HIP(mts::launch_diag_reduce(base, base + total, base + 2 * (size_t)total, total,
                                    base + 3 * (size_t)total, in->stream));
    MTS_HIP(hipMemcpyAsync(r, base + 3 * (size_t)total, sizeof r, hipMemcpyDeviceToHost,
                           in->stream));
    MTS_HIP(hipStreamSynchronize(in->stream));
  }
  // the host report's normalisation: two fictitious points and one jump per crossing
  fp[0] = r[0];
  fp[1] = std::sqrt(r[1] / (2.0 * total));
  jump[0] = r[2];
  jump[1] = std::sqrt(r[3] / (1.0 * total));
  return MTS_OK;
}

int mts_instance_crossing_errors(mts_instance *in, int dir, double t, double *errl,
                                 double *errr, double *jerr, size_t count) {
  if (!in || !errl || !errr || !jerr) return fail(MTS_E_ARG, "null argument");
  if (dir < 0 || dir > 2) return fail(MTS_E_ARG, "direction must be 0, 1 or 2");
  MTS_HIP(hipSetDevice(in->ctx->device));
  int off[4];
  int rc = diag_run(in, t, off);
  if (rc) return rc;
  const size_t n = (size_t)(off[dir + 1] - off[dir]);
  if (count != n) return fail(MTS_E_ARG, "count differs from the direction's crossings");
  if (n == 0) return MTS_OK;
  const size_t total = (size_t)off[3];
  const double *base = in->diag.p;
  MTS_HIP(hipMemcpyAsync(errl, base + off[dir], n * sizeof(double), hipMemcpyDeviceToHost,
                         in->stream));
  MTS_HIP(hipMemcpyAsync(errr, base + total + off[dir], n * sizeof(double),
                         hipMemcpyDeviceToHost, in->stream));
  MTS_HIP(hipMemcpyAsync(jerr, base + 2 * total + off[dir], n * sizeof(double),
                         hipMemcpyDeviceToHost, in->stream));
  MTS_HIP(hipStreamSynchronize(in->stream));
  return MTS_OK;
}

int mts_error_norms_noalloc(mts_instance *in, double t, double *linf, double *l2,
                            int argmax[3]) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  if (!in->ctx->norms_scratch) return fail(MTS_E_DEVICE, "the norms array has no device address");
  if (mts::norms_blocks(in->geom) > mts::kNormsMaxParts)
    return fail(MTS_E_ARG, "grid needs more norm partials than the fixed array holds");
  MTS_HIP(hipSetDevice(in->ctx->device));
  mts::NormsResult *dres = nullptr;
  MTS_HIP(mts::launch_norms_scratch(in->geom, in->field(), in->f32(), std::exp(-t),
                                    in->ctx->norms_scratch, in->stream, &dres));
  mts::NormsResult res;
  MTS_HIP(hipMemcpyAsync(&res, dres, sizeof res, hipMemcpyDeviceToHost, in->stream));
  MTS_HIP(hipStreamSynchronize(in->stream));
  const mts::Grid &g = in->grid;
  const long long ny = g.n[1], nz = g.n[2];
  if (linf) *linf = res.linf;
  if (l2) *l2 = std::sqrt(res.sumsq / (1.0 * g.n[0] * g.n[1] * g.n[2]));
  if (argmax) {
    argmax[0] = (int)(res.argmax / (ny * nz));
    argmax[1] = (int)((res.argmax / nz) % ny);
    argmax[2] = (int)(res.argmax % nz);
  }
  return MTS_OK;
}

int mts_instance_dt(const mts_instance *in, double *dt) {
  if (!in || !dt) return fail(MTS_E_ARG, "null argument");
  *dt = in->dt;
  return MTS_OK;
}

int mts_instance_device_bytes(const mts_instance *in, size_t *bytes) {
  if (!in || !bytes) return fail(MTS_E_ARG, "null argument");
  *bytes = in->device_bytes();
  return MTS_OK;
}

int mts_field_device(mts_instance *in, double **dptr, int *pitch, void **stream) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  if (dptr) *dptr = static_cast<double *>(in->field());  // float storage for an FP32 instance
  if (pitch) *pitch = in->px;
  if (stream) *stream = in->stream;
  return MTS_OK;
}

int mts_profile_enable(mts_instance *in, int on) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  MTS_HIP(hipSetDevice(in->ctx->device));
  MTS_HIP(in->drain());
  in->profile = on != 0;
  return MTS_OK;
}

int mts_profile_reset(mts_instance *in) {
  if (!in) return fail(MTS_E_ARG, "null instance");
  MTS_HIP(hipSetDevice(in->ctx->device));
  MTS_HIP(in->drain());
  for (size_t i = 0; i < in->stage_ms.size(); i++) {
    in->stage_ms[i] = 0.0;
    in->stage_n[i] = 0;
  }
  return MTS_OK;
}

int mts_profile_read(mts_instance *in, mts_stage_time *out, int capacity, int *count) {
  if (!in || !count) return fail(MTS_E_ARG, "null argument");
  MTS_HIP(hipSetDevice(in->ctx->device));
  MTS_HIP(in->drain());
  int n = (int)in->stage_names.size();
  *count = n;
  for (int i = 0; i < n && i < capacity && out; i++) {
    std::snprintf(out[i].name, sizeof(out[i].name), "%s", in->stage_names[i].c_str());
    out[i].launches = in->stage_n[i];
    out[i].total_ms = in->stage_ms[i];
  }
  return MTS_OK;
}

int mts_tdma_batched(mts_ctx *ctx, int n, int nlines, const double *a, const double *b,
                     const double *c, const double *r, double *u) {
  if (!ctx || !a || !b || !c || !r || !u) return fail(MTS_E_ARG, "null argument");
  if (n < 2 || n > 1024 || nlines < 1) return fail(MTS_E_ARG, "bad sizes");
  MTS_HIP(hipSetDevice(ctx->device));
  size_t cnt = (size_t)n * nlines;
  DevBuf<double> da, db, dc, dr, du;
  // everything on the null stream, in order; the last copy is synchronous
  MTS_HIP(da.upload(a, cnt, nullptr));
  MTS_HIP(db.upload(b, cnt, nullptr));
  MTS_HIP(dc.upload(c, cnt, nullptr));
  MTS_HIP(dr.upload(r, cnt, nullptr));
  MTS_HIP(du.alloc(cnt));
  MTS_HIP(mts::launch_tdma(n, nlines, da.p, db.p, dc.p, dr.p, du.p, nullptr));
  MTS_HIP(hipMemcpy(u, du.p, cnt * sizeof(double), hipMemcpyDeviceToHost));
  return MTS_OK;
}

}  // extern "C"
