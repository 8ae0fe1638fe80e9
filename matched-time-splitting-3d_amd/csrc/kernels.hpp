// Host-side declarations of the gfx950 kernels' launchers (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_types.hpp"

namespace mts {

// A second stream and two events: a solve stage split into general and
// regular tiles runs the general tiles on s while the regular ones run on the
// stage's stream (fork: s waits for the work before the stage; join: the
// stage's stream waits for s).  The two tile sets touch disjoint lines.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};

// One line stage along direction dir with the given flags; f32: the fields
// in LineArgs are FP32 (LOD and TS stages only).  side (may be null): where
// the general tiles of a split stage run.
hipError_t launch_line(int dir, int flags, bool f32, const LineArgs &a, hipStream_t s,
                       const SideStream *side = nullptr);

// Per direction, lane-group size and field type (line_kernels.hip: _0 FP64,
// _1 FP32); a.q already filled.
#define MTS_DECL_DL(D, LL)                                                                \
  hipError_t launch_line_dl_##D##_##LL##_0(int flags, bool seq, bool f32, const LineArgs &a, \
                                           hipStream_t s);                                  \
  hipError_t launch_line_dl_##D##_##LL##_1(int flags, bool seq, bool f32, const LineArgs &a, \
                                           hipStream_t s);
MTS_DECL_DL(0, 16) MTS_DECL_DL(0, 32) MTS_DECL_DL(0, 64)
MTS_DECL_DL(1, 16) MTS_DECL_DL(1, 32) MTS_DECL_DL(1, 64)
MTS_DECL_DL(2, 16) MTS_DECL_DL(2, 32) MTS_DECL_DL(2, 64)
#undef MTS_DECL_DL
#define MTS_DECL_MULTI1(D, LL, P)                                                             \
  hipError_t launch_multi_dl_##D##_##LL##_##P(int flags, bool seq, bool f32, int mmax,          \
                                              bool full, bool regular, const LineArgs *hargs,   \
                                              int nargs, const LineArgs *dargs,                 \
                                              const int *dfirst, int nwork, hipStream_t s);
#define MTS_DECL_MULTI(D, LL) MTS_DECL_MULTI1(D, LL, 0) MTS_DECL_MULTI1(D, LL, 1)
MTS_DECL_MULTI(0, 16) MTS_DECL_MULTI(0, 32) MTS_DECL_MULTI(0, 64)
MTS_DECL_MULTI(1, 16) MTS_DECL_MULTI(1, 32) MTS_DECL_MULTI(1, 64)
MTS_DECL_MULTI(2, 16) MTS_DECL_MULTI(2, 32) MTS_DECL_MULTI(2, 64)
#undef MTS_DECL_MULTI
#undef MTS_DECL_MULTI1

// Staging of a multi-instance launch's arguments and work lists: pinned host
// memory (kSlots slots used in turn, each reused once the copy out of it is
// done, so the host runs up to kSlots launches ahead) and one device buffer,
// grown as needed.  Used on one stream.
struct MultiBuf {
  static constexpr int kSlots = 32;
  void *host[kSlots] = {};
  hipEvent_t copied[kSlots] = {};
  size_t host_cap = 0;
  int slot = 0;
  void *dev = nullptr;
  size_t dev_cap = 0;
  MultiBuf() = default;
  MultiBuf(const MultiBuf &) = delete;
  MultiBuf &operator=(const MultiBuf &) = delete;
  ~MultiBuf();
};
// Staging of host bytes for a launch over several instances: parts laid out
// one after the other at 16-byte boundaries (offs), copied into the device
// buffer on s; returns the device base address in *dev.
hipError_t multi_stage(MultiBuf &buf, const void *const *parts, const size_t *sizes, int nparts,
                       size_t *offs, char **dev, hipStream_t s);

// The boundary faces (launch_faces) and Douglas stage 1 (launch_rhs1) of
// several instances, one launch each.
struct FaceArgs {
  DevGeom g;
  PhysConst P;
  void *out;
  double e, cf;
};
struct Rhs1Multi {
  Rhs1Args p;
  SpecArgs sp;
  unsigned npairs, nper, hp;
  double inv_hp;
};
hipError_t launch_faces_multi(const FaceArgs *args, int count, bool f32, MultiBuf &buf,
                              hipStream_t s);
hipError_t launch_rhs1_multi(const Rhs1Args *r1, const SpecArgs *sp, int count, bool pairs_too,
                             bool f32, MultiBuf &buf, hipStream_t s);

// One solve stage (dir, flags) of count instances in as few launches as
// their kernels allow: the general and the regular tiles of every instance
// whose split, register bound, full split and sequential lines match go into
// one launch over all their tiles (line_kernel_multi / line_kernel_reg_multi).
// args[i] are the instances' stage arguments as launch_line takes them (a
// split solve stage).  Same arithmetic as launch_line per instance.
hipError_t launch_line_multi(int dir, int flags, bool f32, const LineArgs *args, int count,
                             MultiBuf &buf, hipStream_t s);

// Douglas stage 1 right-hand side (see device_core.hpp): the regular
// stencil on every node but the MIB-affected ones (classification +-2), and
// those from sp (spec_node), in one launch; pairs_too false: the
// MIB-affected nodes only (the x sweep computes the rest, kFlagPre1).
hipError_t launch_rhs1(const Rhs1Args &a, const SpecArgs &sp, bool pairs_too, bool f32,
                       hipStream_t s);

// Boundary faces: out = exact outer solution at e (+ cf * f(e) when cf != 0);
// f32: out is an FP32 field.
hipError_t launch_faces(const DevGeom &g, const PhysConst &P, void *out, bool f32, double e,
                        double cf, hipStream_t s);

// Exact solution at e = exp(-t0) on every node (Initialization).
hipError_t launch_init(const DevGeom &g, void *u, bool f32, double e, hipStream_t s);

// Error norms against the exact solution: per-block partial (max, ref index,
// sum of squares); nblocks written to *nparts.  Accumulated in FP64.
hipError_t launch_norms(const DevGeom &g, const void *u, bool f32, double e, double *part_max,
                        long long *part_idx, double *part_sq, int *nparts, hipStream_t s);
int norms_blocks(const DevGeom &g);

// Error norms with the partials and their reduction in a fixed device array
// of the library (kNormsScratchBytes per GPU): no device allocation.  The
// reduced record is left in that array at *result (a device address).  Calls
// on one GPU share the array: calls that overlap on a GPU race.
constexpr int kNormsMaxParts = 2048;  // partials of one norms launch, at most
constexpr size_t kNormsScratchBytes = (size_t)kNormsMaxParts * 24 + 256;
struct NormsResult {
  double linf, sumsq;
  long long argmax;  // reference-order node index of the first maximum
};
// The array's device address on device (once per device; the device must be
// current on the first call).
hipError_t norms_scratch_ptr(int device, void **ptr);
hipError_t launch_norms_scratch(const DevGeom &g, const void *u, bool f32, double e,
                                void *scratch_ptr, hipStream_t s, NormsResult **result);

// Interface diagnostics of direction dir (per crossing, into a.el / a.er /
// a.je), and their reduction over n crossings into out[0..3] = max and sum of
// squares of the fictitious-point errors, then of the jump errors.
hipError_t launch_diagnostics(int dir, const DiagArgs &a, hipStream_t s);
hipError_t launch_diag_reduce(const double *el, const double *er, const double *je, int n,
                              double *out, hipStream_t s);

// Batched tridiagonal solve: nlines independent systems of length n, arrays
// line-major (line l, row i at l*n + i).  a[l*n] and c[l*n + n-1] are ignored.
hipError_t launch_tdma(int n, int nlines, const double *a, const double *b, const double *c,
                       const double *r, double *u, hipStream_t s);

}  // namespace mts
