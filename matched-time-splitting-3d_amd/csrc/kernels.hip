// gfx950 kernels of the Matched-TS per-step path.
//
// Every implicit stage of Douglas-ADI, LOD-IE/CN and Trapezoidal Splitting is
// a set of independent tridiagonal line systems (the reference solves each
// with TDMA, e.g. temporal/douglas_adi.cc:1784-1807).  Here one kernel
// launch does one stage along one direction for all lines:
//
//   * a workgroup owns a tile of 16 neighbouring lines, stages the stage's
//     input field(s) and the node classification for the whole tile in LDS
//     with coalesced loads (y/z lines: 16 consecutive x = one 128-byte row
//     segment per line row; x lines: contiguous rows),
//   * one 64-lane wavefront per line builds the right-hand side of its rows
//     (3-point beta-weighted stencil, analytic source and boundary values
//     from per-axis sin/cos tables, MIB fictitious-point rows from the sparse
//     block table), then solves the line with a partitioned Thomas / PCR
//     hybrid: each lane eliminates its own segment of rows in registers,
//     the 2-unknown-per-lane reduced system is folded to one unknown per lane
//     and solved by parallel cyclic reduction across the wavefront with lane
//     shuffles, and each lane back-substitutes its rows,
//   * the tile is written back with coalesced stores.
//
// All arithmetic mirrors the reference's operation order (fp contraction is
// disabled for this file) except the tridiagonal elimination order.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "device_core.hpp"
#include "kernels.hpp"

namespace mts {

namespace {

constexpr int kBlock = 256;   // threads per workgroup (4 wavefronts)

// -------------------------------------------------- wavefront line solver
// Lane k holds rows [s_k, s_k + m) of one tridiagonal system
//   a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i
// (m >= 2 on active lanes, m == 0 on idle lanes; the first row of the system
// has a = 0 and the last c = 0).  On return the segment's first and last
// unknowns are in xs and xe and the rows in between in d (see solved_row).
template <int MMAX>
__device__ __forceinline__ void wave_solve(double (&a)[MMAX], double (&b)[MMAX],
                                           double (&c)[MMAX], double (&d)[MMAX], int m,
                                           int lane, double &xs, double &xe) {
  // reduced rows:  A1 x_{s-1} + B1 x_s + C1 x_e = D1   (first row)
  //                A2 x_s + B2 x_e + C2 x_{e+1} = D2   (last row)
  double A1 = 0.0, B1 = 1.0, C1 = 0.0, D1 = 0.0;
  double A2 = 0.0, B2 = 1.0, C2 = 0.0, D2 = 0.0;
  if (m >= 2) {
    // downward elimination from row 1 keeping x_s as a parameter:
    // alpha_i x_s + x_i + gamma_i x_{i+1} = delta_i  (stored in a, c, d)
    {
      double r = 1.0 / b[1];
      a[1] = a[1] * r;
      c[1] = c[1] * r;
      d[1] = d[1] * r;
      A2 = a[1];
      C2 = c[1];
      D2 = d[1];
    }
#pragma unroll
    for (int i = 2; i < MMAX; i++) {
      if (i < m) {
        double r = 1.0 / (b[i] - a[i] * c[i - 1]);
        d[i] = (d[i] - a[i] * d[i - 1]) * r;
        a[i] = -a[i] * a[i - 1] * r;
        c[i] = c[i] * r;
        // the last row reached is the segment's last row
        A2 = a[i];
        C2 = c[i];
        D2 = d[i];
      }
    }
    // upward elimination down to row 1, expressing x_i through x_s and x_e
#pragma unroll
    for (int i = MMAX - 3; i >= 1; i--) {
      if (i <= m - 3) {
        d[i] = d[i] - c[i] * d[i + 1];
        a[i] = a[i] - c[i] * a[i + 1];
        c[i] = -c[i] * c[i + 1];
      }
    }
    if (m == 2) {
      A1 = a[0];
      B1 = b[0];
      C1 = c[0];
      D1 = d[0];
    } else {
      A1 = a[0];
      B1 = b[0] - c[0] * a[1];
      C1 = -c[0] * c[1];
      D1 = d[0] - c[0] * d[1];
    }
  }
  // fold x_s out: one equation per lane in the segment-end unknowns x_e
  double A1n = __shfl_down(A1, 1, 64);
  double B1n = __shfl_down(B1, 1, 64);
  double C1n = __shfl_down(C1, 1, 64);
  double D1n = __shfl_down(D1, 1, 64);
  double ea = 0.0, eb = 1.0, ec = 0.0, ed = 0.0;
  if (m >= 2) {
    if (lane == 63) {
      A1n = 0.0;
      B1n = 1.0;
      C1n = 0.0;
      D1n = 0.0;
    }
    double i1 = 1.0 / B1, in = 1.0 / B1n;
    ea = -A2 * A1 * i1;
    eb = B2 - A2 * C1 * i1 - C2 * A1n * in;
    ec = -C2 * C1n * in;
    ed = D2 - A2 * D1 * i1 - C2 * D1n * in;
  }
  // parallel cyclic reduction across the wavefront
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) {
    double am = __shfl_up(ea, s, 64), bm = __shfl_up(eb, s, 64);
    double cm = __shfl_up(ec, s, 64), dm = __shfl_up(ed, s, 64);
    double ap = __shfl_down(ea, s, 64), bp = __shfl_down(eb, s, 64);
    double cp = __shfl_down(ec, s, 64), dp = __shfl_down(ed, s, 64);
    if (lane < s) {
      am = 0.0; bm = 1.0; cm = 0.0; dm = 0.0;
    }
    if (lane + s >= 64) {
      ap = 0.0; bp = 1.0; cp = 0.0; dp = 0.0;
    }
    double k1 = ea / bm, k2 = ec / bp;
    double na = -am * k1;
    double nc = -cp * k2;
    double nb = eb - cm * k1 - ap * k2;
    double nd = ed - dm * k1 - dp * k2;
    ea = na;
    eb = nb;
    ec = nc;
    ed = nd;
  }
  xe = ed / eb;
  double xprev = __shfl_up(xe, 1, 64);
  if (lane == 0) xprev = 0.0;
  xs = xe;
  if (m >= 2) {
    xs = (D1 - A1 * xprev - C1 * xe) / B1;
#pragma unroll
    for (int i = 1; i < MMAX - 1; i++) {
      if (i < m - 1) d[i] = d[i] - a[i] * xs - c[i] * xe;
    }
  }
}

// value of local row r after wave_solve
template <int MMAX>
__device__ __forceinline__ double solved_row(const double (&d)[MMAX], int r, int m, double xs,
                                             double xe) {
  return r == 0 ? xs : (r == m - 1 ? xe : d[r]);
}

// ------------------------------------------------ Douglas stage-1 RHS
// One launch: the first nper * 8 workgroups compute R on node pairs (one
// thread per pair of neighbouring x nodes, pairs in x-fastest order with
// px / 2 per x row, 256 per workgroup), the rest on the MIB-affected nodes
// (spec_node), which the pair threads leave alone.  The hardware deals
// workgroups to the 8 XCDs round-robin; pair workgroup b is given pair block
// (b mod 8) nper + b / 8, so each XCD walks one contiguous run of z-planes
// and the y and z neighbours a block reads were fetched into that XCD's L2 by
// the blocks just before it.
// The instance of workgroup b of a launch over several instances, whose
// first workgroups are first[0..count) (ascending, first[0] = 0).
__device__ __forceinline__ int multi_instance(const int *first, int count, unsigned b) {
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if ((unsigned)first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
constexpr int kXcds = 8;
template <typename F>
__device__ __forceinline__ void rhs1_block(const Rhs1Args &p, const SpecArgs &sp, unsigned npairs,
                                           unsigned nper, unsigned hp, double inv_hp,
                                           unsigned blk) {
  if (blk >= nper * kXcds) {
    const int t = (int)(blk - nper * kXcds) * 256 + (int)threadIdx.x;
    if (t < sp.nspec) spec_node<F>(sp, t);
    return;
  }
  const unsigned b = (blk % kXcds) * nper + blk / kXcds;
  const unsigned t = b * 256u + threadIdx.x;
  if (t >= npairs) return;
  // t / hp through a double reciprocal, corrected to the exact quotient
  unsigned row = (unsigned)((double)t * inv_hp);
  if (row * hp > t) row--;
  else if ((row + 1) * hp <= t) row++;
  rhs1_pair<F>(p, row, t - row * hp);
}
template <typename F>
__global__ __launch_bounds__(256) void rhs1_kernel(Rhs1Args p, SpecArgs sp, unsigned npairs,
                                                   unsigned nper, unsigned hp, double inv_hp) {
  rhs1_block<F>(p, sp, npairs, nper, hp, inv_hp, blockIdx.x);
}
// Douglas stage 1 of several instances in one launch (each instance's
// workgroups as rhs1_kernel's).
template <typename F>
__global__ __launch_bounds__(256) void rhs1_multi_kernel(const Rhs1Multi *args, const int *first,
                                                         int count) {
  const int i = multi_instance(first, count, blockIdx.x);
  const Rhs1Multi &a = args[i];
  rhs1_block<F>(a.p, a.sp, a.npairs, a.nper, a.hp, a.inv_hp, blockIdx.x - (unsigned)first[i]);
}

// ------------------------------------------------------- boundary faces
template <typename R>
__device__ __forceinline__ void face_node(const DevGeom &g, const PhysConst &P, R *out, double e,
                                          double cf, int t) {
  // (32-bit index arithmetic: the six faces hold at most 6 * 1024^2 nodes,
  // and 64-bit division is a long software sequence on the GPU)
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int fz = nx * ny, fy = nx * nz, fx = ny * nz;
  int ix, iy, iz;
  if (t < 2 * fz) {
    iz = t < fz ? 0 : nz - 1;
    const int q = t < fz ? t : t - fz;
    iy = q / nx;
    ix = q - iy * nx;
  } else if ((t -= 2 * fz) < 2 * fy) {
    iy = t < fy ? 0 : ny - 1;
    const int q = t < fy ? t : t - fy;
    iz = q / nx;
    ix = q - iz * nx;
  } else if ((t -= 2 * fy) < 2 * fx) {
    ix = t < fx ? 0 : nx - 1;
    const int q = t < fx ? t : t - fx;
    iz = q / ny;
    iy = q - iz * ny;
  } else {
    return;
  }
  out[node_index(g, ix, iy, iz)] = (R)face_value(P, g, e, cf, ix, iy, iz);
}
template <typename R>
__global__ __launch_bounds__(256) void faces_kernel(DevGeom g, PhysConst P, R *out, double e,
                                                    double cf) {
  face_node<R>(g, P, out, e, cf, (int)(blockIdx.x * blockDim.x + threadIdx.x));
}
// The boundary faces of several instances in one launch.
template <typename R>
__global__ __launch_bounds__(256) void faces_multi_kernel(const FaceArgs *args, const int *first,
                                                          int count) {
  const int i = multi_instance(first, count, blockIdx.x);
  const FaceArgs &a = args[i];
  face_node<R>(a.g, a.P, static_cast<R *>(a.out), a.e, a.cf,
               (int)((blockIdx.x - (unsigned)first[i]) * blockDim.x + threadIdx.x));
}

// ------------------------------------------------------- initialization
template <typename R>
__global__ __launch_bounds__(256) void init_kernel(DevGeom g, R *u, double e) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y, iz = blockIdx.z;
  if (ix >= g.n[0]) return;
  const size_t c = node_index(g, ix, iy, iz);
  u[c] = (R)(g.ind[c] > 0 ? outer_u(g, e, ix, iy, iz) : inner_u(g, e, ix, iy, iz));
}

// ------------------------------------------------------------ error norms
// Reference order index (ix outer, iz inner, douglas_adi.cc:1736-1760) so the
// first maximum in that order wins ties.
template <typename R>
__global__ __launch_bounds__(256) void norms_kernel(DevGeom g, const R *u, double e,
                                                    double *pmax, long long *pidx,
                                                    double *psq) {
  __shared__ double smax[256];
  __shared__ long long sidx[256];
  __shared__ double ssq[256];
  const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const long long total = (long long)nx * ny * nz;
  double mx = -1.0, sq = 0.0;
  long long mi = total;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int ix = (int)(t % nx);
    long long r = t / nx;
    int iy = (int)(r % ny);
    int iz = (int)(r / ny);
    size_t c = node_index(g, ix, iy, iz);
    double ex = g.ind[c] > 0 ? outer_u(g, e, ix, iy, iz) : inner_u(g, e, ix, iy, iz);
    double er = fabs((double)u[c] - ex);
    long long ri = ((long long)ix * ny + iy) * nz + iz;
    if (er > mx || (er == mx && ri < mi)) {
      mx = er;
      mi = ri;
    }
    sq += er * er;
  }
  smax[threadIdx.x] = mx;
  sidx[threadIdx.x] = mi;
  ssq[threadIdx.x] = sq;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double om = smax[threadIdx.x + w];
      long long oi = sidx[threadIdx.x + w];
      if (om > smax[threadIdx.x] || (om == smax[threadIdx.x] && oi < sidx[threadIdx.x])) {
        smax[threadIdx.x] = om;
        sidx[threadIdx.x] = oi;
      }
      ssq[threadIdx.x] += ssq[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmax[blockIdx.x] = smax[0];
    pidx[blockIdx.x] = sidx[0];
    psq[blockIdx.x] = ssq[0];
  }
}

// The partials reduced in the order of the host loop of mts_error_norms
// (first maximum by reference index, squares summed part 0 first), so the
// result is bitwise the same.  One thread: at most kNormsMaxParts partials.
__global__ __launch_bounds__(64) void norms_final_kernel(const double *pmax,
                                                         const long long *pidx,
                                                         const double *psq, int nparts,
                                                         NormsResult *out) {
  if (threadIdx.x != 0) return;
  double mx = -1.0, sq = 0.0;
  long long mi = 0;
  for (int i = 0; i < nparts; i++) {
    if (pmax[i] > mx || (pmax[i] == mx && pidx[i] < mi)) {
      mx = pmax[i];
      mi = pidx[i];
    }
    sq += psq[i];
  }
  out->linf = mx;
  out->sumsq = sq;
  out->argmax = mi;
}

// ------------------------------------------------ interface diagnostics
// One thread per line of the direction (lines without blocks return at once).
template <int DIR>
__global__ __launch_bounds__(256) void diag_kernel(DiagArgs a) {
  const int lid = (int)(blockIdx.x * 256 + threadIdx.x);
  if (lid >= a.nj * a.nk) return;
  diag_line<DIR>(a, lid);
}
// One workgroup: max |x| and the sum of squares of the fictitious-point errors
// (el, er: n each) and of the jump errors (je), in a fixed order (each thread
// a fixed stride, then a tree), into out[0..3] = fmax, fsum, jmax, jsum.
// NaNs drop out of the maxima as in the host loop (a NaN is never larger).
__global__ __launch_bounds__(1024) void diag_reduce_kernel(const double *el, const double *er,
                                                           const double *je, int n, double *out) {
  __shared__ double sh[4][1024];
  double fm = 0.0, fs = 0.0, jm = 0.0, js = 0.0;
  for (int i = (int)threadIdx.x; i < n; i += 1024) {
    const double l = el[i], r = er[i], q = je[i];
    fs += l * l;
    fs += r * r;
    if (fabs(l) > fm) fm = fabs(l);
    if (fabs(r) > fm) fm = fabs(r);
    js += q * q;
    if (fabs(q) > jm) jm = fabs(q);
  }
  const int t = (int)threadIdx.x;
  sh[0][t] = fm;
  sh[1][t] = fs;
  sh[2][t] = jm;
  sh[3][t] = js;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) {
      if (sh[0][t + w] > sh[0][t]) sh[0][t] = sh[0][t + w];
      sh[1][t] += sh[1][t + w];
      if (sh[2][t + w] > sh[2][t]) sh[2][t] = sh[2][t + w];
      sh[3][t] += sh[3][t + w];
    }
    __syncthreads();
  }
  if (t < 4) out[t] = sh[t][0];
}

// ----------------------------------------------------- batched tridiagonal
template <int MMAX>
__global__ __launch_bounds__(256) void tdma_kernel(int n, int nlines, const double *A,
                                                   const double *B, const double *C,
                                                   const double *R, double *U) {
  const int lane = threadIdx.x & 63;
  const int line = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (line >= nlines) return;  // uniform per wavefront
  int s, m;
  segment(n, lane, s, m);
  const size_t base = (size_t)line * n;
  double a[MMAX], b[MMAX], c[MMAX], d[MMAX];
#pragma unroll
  for (int r = 0; r < MMAX; r++) {
    a[r] = 0.0;
    b[r] = 1.0;
    c[r] = 0.0;
    d[r] = 0.0;
    if (r < m) {
      const int i = s + r;
      a[r] = i == 0 ? 0.0 : A[base + i];
      b[r] = B[base + i];
      c[r] = i == n - 1 ? 0.0 : C[base + i];
      d[r] = R[base + i];
    }
  }
  double xs, xe;
  wave_solve<MMAX>(a, b, c, d, m, lane, xs, xe);
#pragma unroll
  for (int r = 0; r < MMAX; r++)
    if (r < m) U[base + s + r] = solved_row(d, r, m, xs, xe);
}

// Device memory of the allocation-free error norms (mts_error_norms_noalloc):
// the per-block partials, then the reduced result.  Each GPU holds its own
// copy from the library's load until the process ends.  In host code the name
// is only a handle for hipGetSymbolAddress; kernels get the address it
// returns (norms_scratch_ptr).
constexpr size_t kScratchPmax = 0;
constexpr size_t kScratchPsq = kScratchPmax + kNormsMaxParts * sizeof(double);
constexpr size_t kScratchPidx = kScratchPsq + kNormsMaxParts * sizeof(double);
constexpr size_t kScratchResult = kScratchPidx + kNormsMaxParts * sizeof(long long);
constexpr size_t kScratchBytes = kScratchResult + 256;
static_assert(kScratchBytes == kNormsScratchBytes, "scratch layout");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

}  // namespace

// ------------------------------------------------------ allocation-free norms
namespace {
std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // per device, filled once
}  // namespace

hipError_t norms_scratch_ptr(int device, void **ptr) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0) return hipErrorInvalidDevice;
  if ((size_t)device < g_scratch_ptr.size() && g_scratch_ptr[(size_t)device]) {
    *ptr = g_scratch_ptr[(size_t)device];
    return hipSuccess;
  }
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (g_scratch_ptr.size() <= (size_t)device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  g_scratch_ptr[(size_t)device] = p;
  *ptr = p;
  return hipSuccess;
}

hipError_t launch_norms_scratch(const DevGeom &g, const void *u, bool f32, double e,
                                void *scratch_ptr, hipStream_t s, NormsResult **result) {
  if (!scratch_ptr) return hipErrorInvalidValue;
  const int nb = norms_blocks(g);
  if (nb < 1 || nb > kNormsMaxParts) return hipErrorInvalidValue;
  unsigned char *base = static_cast<unsigned char *>(scratch_ptr);
  double *pmax = reinterpret_cast<double *>(base + kScratchPmax);
  double *psq = reinterpret_cast<double *>(base + kScratchPsq);
  long long *pidx = reinterpret_cast<long long *>(base + kScratchPidx);
  NormsResult *res = reinterpret_cast<NormsResult *>(base + kScratchResult);
  if (f32)
    hipLaunchKernelGGL(norms_kernel<float>, dim3(nb), dim3(256), 0, s, g,
                       static_cast<const float *>(u), e, pmax, pidx, psq);
  else
    hipLaunchKernelGGL(norms_kernel<double>, dim3(nb), dim3(256), 0, s, g,
                       static_cast<const double *>(u), e, pmax, pidx, psq);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(norms_final_kernel, dim3(1), dim3(64), 0, s, pmax, pidx, psq, nb, res);
  *result = res;
  return hipGetLastError();
}

// --------------------------------------------------------------- launchers
hipError_t launch_line(int dir, int flags, bool f32, const LineArgs &args, hipStream_t s,
                       const SideStream *side) {
  auto one = [&](const LineArgs &in, int min_lanes, hipStream_t s) -> hipError_t {
    LineArgs a = in;
    if (!make_line_part(a.g.n[dir], a.q, min_lanes)) return hipErrorInvalidValue;
    if (tile_lines(a.q.L) * a.g.n[dir] > kTileElsPerThread * kBlockThreads) return hipErrorInvalidValue;
    const bool seq = (flags & kFlagSolve) && a.mib.seq != nullptr;
#define MTS_DL(D, LL) \
  if (dir == D && a.q.L == LL)                                                           \
    return f32 ? launch_line_dl_##D##_##LL##_1(flags, seq, f32, a, s)                    \
               : launch_line_dl_##D##_##LL##_0(flags, seq, f32, a, s);
    MTS_DL(0, 16) MTS_DL(0, 32) MTS_DL(0, 64)
    MTS_DL(1, 16) MTS_DL(1, 32) MTS_DL(1, 64)
    MTS_DL(2, 16) MTS_DL(2, 32) MTS_DL(2, 64)
#undef MTS_DL
    return hipErrorInvalidValue;
  };
  if ((flags & kFlagSolve) && args.split) {
    // regular tiles at the stage's split, general ones at gen_lanes lanes
    LineArgs ar = args, ag = args;
    ar.ntiles[1] = 0;
    ag.ntiles[0] = 0;
    const int gl = args.gen_lanes > args.min_lanes ? args.gen_lanes : args.min_lanes;
    if (side && side->s && ag.ntiles[1] > 0 && ar.ntiles[0] > 0) {
      // both kinds at once: the few, long general tiles first, on the side
      // stream, and the regular tiles beside them
      hipError_t e = hipEventRecord(side->fork, s);
      if (e == hipSuccess) e = hipStreamWaitEvent(side->s, side->fork, 0);
      if (e == hipSuccess) e = one(ag, gl, side->s);
      if (e == hipSuccess) e = one(ar, args.min_lanes, s);
      if (e == hipSuccess) e = hipEventRecord(side->join, side->s);
      if (e == hipSuccess) e = hipStreamWaitEvent(s, side->join, 0);
      return e;
    }
    hipError_t e = hipSuccess;
    if (ag.ntiles[1] > 0) e = one(ag, gl, s);
    if (e == hipSuccess && ar.ntiles[0] > 0) e = one(ar, args.min_lanes, s);
    return e;
  }
  return one(args, args.min_lanes, s);
}

MultiBuf::~MultiBuf() {
  for (int i = 0; i < kSlots; i++) {
    if (copied[i]) (void)hipEventSynchronize(copied[i]);
    if (copied[i]) (void)hipEventDestroy(copied[i]);
    if (host[i]) (void)hipHostFree(host[i]);
  }
  if (dev) (void)hipFree(dev);
}

hipError_t multi_stage(MultiBuf &buf, const void *const *parts, const size_t *sizes, int nparts,
                       size_t *offs, char **dev, hipStream_t s) {
  size_t bytes = 0;
  for (int i = 0; i < nparts; i++) {
    offs[i] = bytes;
    bytes = (bytes + sizes[i] + 15) / 16 * 16;
  }
  hipError_t e = hipSuccess;
  if (bytes > buf.host_cap || bytes > buf.dev_cap) {
    e = hipStreamSynchronize(s);  // (nothing in flight reads the old buffers)
    if (e != hipSuccess) return e;
    const size_t cap = bytes * 2;
    for (int i = 0; i < MultiBuf::kSlots; i++) {
      if (buf.copied[i]) {
        e = hipEventSynchronize(buf.copied[i]);
        if (e != hipSuccess) return e;
      }
      if (buf.host[i]) (void)hipHostFree(buf.host[i]);
      buf.host[i] = nullptr;
      e = hipHostMalloc(&buf.host[i], cap, hipHostMallocDefault);
      if (e != hipSuccess) return e;
      if (!buf.copied[i]) {
        e = hipEventCreateWithFlags(&buf.copied[i], hipEventDisableTiming);
        if (e != hipSuccess) return e;
      }
    }
    if (buf.dev) (void)hipFree(buf.dev);
    buf.dev = nullptr;
    e = hipMalloc(&buf.dev, cap);
    if (e != hipSuccess) return e;
    buf.host_cap = buf.dev_cap = cap;
  }
  // the slot's last copy has left it
  const int sl = buf.slot;
  buf.slot = (buf.slot + 1) % MultiBuf::kSlots;
  e = hipEventSynchronize(buf.copied[sl]);
  if (e != hipSuccess) return e;
  char *h = static_cast<char *>(buf.host[sl]);
  for (int i = 0; i < nparts; i++)
    if (sizes[i]) std::memcpy(h + offs[i], parts[i], sizes[i]);
  e = hipMemcpyAsync(buf.dev, h, bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipEventRecord(buf.copied[sl], s);
  *dev = static_cast<char *>(buf.dev);
  return e;
}

hipError_t launch_faces_multi(const FaceArgs *args, int count, bool f32, MultiBuf &buf,
                              hipStream_t s) {
  if (count < 1) return hipSuccess;
  std::vector<int> first((size_t)count);
  long long blocks = 0;
  for (int i = 0; i < count; i++) {
    const DevGeom &g = args[i].g;
    first[(size_t)i] = (int)blocks;
    const long long total = 2LL * ((long long)g.n[0] * g.n[1] + (long long)g.n[0] * g.n[2] +
                                   (long long)g.n[1] * g.n[2]);
    blocks += (total + 255) / 256;
  }
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const void *parts[2] = {args, first.data()};
  const size_t sizes[2] = {(size_t)count * sizeof(FaceArgs), first.size() * sizeof(int)};
  size_t offs[2];
  char *d = nullptr;
  hipError_t e = multi_stage(buf, parts, sizes, 2, offs, &d, s);
  if (e != hipSuccess) return e;
  const FaceArgs *da = reinterpret_cast<const FaceArgs *>(d + offs[0]);
  const int *df = reinterpret_cast<const int *>(d + offs[1]);
  if (f32)
    hipLaunchKernelGGL(faces_multi_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, s, da, df,
                       count);
  else
    hipLaunchKernelGGL(faces_multi_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, s, da, df,
                       count);
  return hipGetLastError();
}

hipError_t launch_rhs1_multi(const Rhs1Args *r1, const SpecArgs *sp, int count, bool pairs_too,
                             bool f32, MultiBuf &buf, hipStream_t s) {
  if (count < 1) return hipSuccess;
  std::vector<Rhs1Multi> m((size_t)count);
  std::vector<int> first((size_t)count);
  long long blocks = 0;
  for (int i = 0; i < count; i++) {
    const Rhs1Args &a = r1[i];
    if (a.g.px % 2 != 0 || sp[i].nspec < 0) return hipErrorInvalidValue;
    const unsigned hp = (unsigned)(a.g.px / 2);
    const size_t pairs = pairs_too ? (size_t)hp * a.g.n[1] * a.g.n[2] : 0;
    if (pairs > 0xffffffffu - 256u * kXcds) return hipErrorInvalidValue;
    Rhs1Multi &q = m[(size_t)i];
    q.p = a;
    q.sp = sp[i];
    q.npairs = (unsigned)pairs;
    q.nper = (unsigned)(((pairs + 255) / 256 + kXcds - 1) / kXcds);
    q.hp = hp;
    q.inv_hp = 1.0 / hp;
    first[(size_t)i] = (int)blocks;
    blocks += (long long)q.nper * kXcds + (sp[i].nspec + 255) / 256;
  }
  if (blocks == 0) return hipSuccess;
  if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
  const void *parts[2] = {m.data(), first.data()};
  const size_t sizes[2] = {m.size() * sizeof(Rhs1Multi), first.size() * sizeof(int)};
  size_t offs[2];
  char *d = nullptr;
  hipError_t e = multi_stage(buf, parts, sizes, 2, offs, &d, s);
  if (e != hipSuccess) return e;
  const Rhs1Multi *da = reinterpret_cast<const Rhs1Multi *>(d + offs[0]);
  const int *df = reinterpret_cast<const int *>(d + offs[1]);
  if (f32)
    hipLaunchKernelGGL(rhs1_multi_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, s, da, df,
                       count);
  else
    hipLaunchKernelGGL(rhs1_multi_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, s, da, df,
                       count);
  return hipGetLastError();
}

hipError_t launch_line_multi(int dir, int flags, bool f32, const LineArgs *args, int count,
                             MultiBuf &buf, hipStream_t s) {
  if (!(flags & kFlagSolve) || dir < 0 || dir > 2) return hipErrorInvalidValue;
  struct Key {
    int L, mmax, full, seq, regular;
    bool operator==(const Key &o) const {
      return L == o.L && mmax == o.mmax && full == o.full && seq == o.seq && regular == o.regular;
    }
  };
  struct Group {
    Key k;
    std::vector<LineArgs> args;
    std::vector<int> first;  // each slot's first workgroup
    int blocks = 0;
  };
  std::vector<Group> groups;
  for (int i = 0; i < count; i++) {
    const LineArgs &in = args[i];
    if (!in.split) return hipErrorInvalidValue;
    for (int regular = 0; regular < 2; regular++) {
      const int nt = in.ntiles[regular ? 0 : 1];
      if (nt <= 0) continue;
      LineArgs a = in;
      a.ntiles[regular ? 1 : 0] = 0;
      const int lanes = regular ? in.min_lanes
                                : (in.gen_lanes > in.min_lanes ? in.gen_lanes : in.min_lanes);
      if (!make_line_part(a.g.n[dir], a.q, lanes)) return hipErrorInvalidValue;
      if (tile_lines(a.q.L) * a.g.n[dir] > kTileElsPerThread * kBlockThreads)
        return hipErrorInvalidValue;
      Key k{a.q.L, a.q.mmax, a.q.m == a.q.mmax ? 1 : 0, a.mib.seq != nullptr ? 1 : 0, regular};
      size_t gi = 0;
      while (gi < groups.size() && !(groups[gi].k == k)) gi++;
      if (gi == groups.size()) groups.push_back(Group{k, {}, {}, 0});
      Group &g = groups[gi];
      g.args.push_back(a);
      g.first.push_back(g.blocks);
      g.blocks += nt;
    }
  }
  if (groups.empty()) return hipSuccess;
  // one staging image: per group its arguments, then its work entries
  std::vector<const void *> parts;
  std::vector<size_t> sizes;
  for (const Group &g : groups) {
    parts.push_back(g.args.data());
    sizes.push_back(g.args.size() * sizeof(LineArgs));
    parts.push_back(g.first.data());
    sizes.push_back(g.first.size() * sizeof(int));
  }
  std::vector<size_t> offs(parts.size());
  char *d = nullptr;
  hipError_t e = multi_stage(buf, parts.data(), sizes.data(), (int)parts.size(), offs.data(), &d, s);
  if (e != hipSuccess) return e;
  for (size_t gi = 0; gi < groups.size() && e == hipSuccess; gi++) {
    const Group &g = groups[gi];
    const LineArgs *da = reinterpret_cast<const LineArgs *>(d + offs[2 * gi]);
    const int *dw = reinterpret_cast<const int *>(d + offs[2 * gi + 1]);
    const int na = (int)g.args.size(), nw = g.blocks;
    const bool full = g.k.full != 0, seq = g.k.seq != 0, reg = g.k.regular != 0;
#define MTS_ML(D, LL)                                                                          \
  if (dir == D && g.k.L == LL) {                                                               \
    e = f32 ? launch_multi_dl_##D##_##LL##_1(flags, seq, f32, g.k.mmax, full, reg,            \
                                             g.args.data(), na, da, dw, nw, s)                 \
            : launch_multi_dl_##D##_##LL##_0(flags, seq, f32, g.k.mmax, full, reg,            \
                                             g.args.data(), na, da, dw, nw, s);                \
    continue;                                                                                  \
  }
    MTS_ML(0, 16) MTS_ML(0, 32) MTS_ML(0, 64)
    MTS_ML(1, 16) MTS_ML(1, 32) MTS_ML(1, 64)
    MTS_ML(2, 16) MTS_ML(2, 32) MTS_ML(2, 64)
#undef MTS_ML
    e = hipErrorInvalidValue;
  }
  return e;
}

hipError_t launch_rhs1(const Rhs1Args &a, const SpecArgs &sp, bool pairs_too, bool f32,
                       hipStream_t s) {
  if (a.g.px % 2 != 0 || sp.nspec < 0) return hipErrorInvalidValue;
  const unsigned hp = (unsigned)(a.g.px / 2);
  const size_t pairs = pairs_too ? (size_t)hp * a.g.n[1] * a.g.n[2] : 0;
  if (pairs > 0xffffffffu - 256u * kXcds) return hipErrorInvalidValue;
  const unsigned nper = (unsigned)(((pairs + 255) / 256 + kXcds - 1) / kXcds);
  const unsigned nspec_blocks = (unsigned)((sp.nspec + 255) / 256);
  if (nper * kXcds + nspec_blocks == 0) return hipSuccess;
  const dim3 grid(nper * kXcds + nspec_blocks);
  if (f32)
    hipLaunchKernelGGL(rhs1_kernel<float>, grid, dim3(256), 0, s, a, sp, (unsigned)pairs, nper,
                       hp, 1.0 / hp);
  else
    hipLaunchKernelGGL(rhs1_kernel<double>, grid, dim3(256), 0, s, a, sp, (unsigned)pairs, nper,
                       hp, 1.0 / hp);
  return hipGetLastError();
}

hipError_t launch_faces(const DevGeom &g, const PhysConst &P, void *out, bool f32, double e,
                        double cf, hipStream_t s) {
  long long total = 2LL * ((long long)g.n[0] * g.n[1] + (long long)g.n[0] * g.n[2] +
                           (long long)g.n[1] * g.n[2]);
  const dim3 grid((unsigned)((total + 255) / 256));
  if (f32)
    hipLaunchKernelGGL(faces_kernel<float>, grid, dim3(256), 0, s, g, P, static_cast<float *>(out),
                       e, cf);
  else
    hipLaunchKernelGGL(faces_kernel<double>, grid, dim3(256), 0, s, g, P,
                       static_cast<double *>(out), e, cf);
  return hipGetLastError();
}

hipError_t launch_init(const DevGeom &g, void *u, bool f32, double e, hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  if (f32)
    hipLaunchKernelGGL(init_kernel<float>, grid, dim3(256), 0, s, g, static_cast<float *>(u), e);
  else
    hipLaunchKernelGGL(init_kernel<double>, grid, dim3(256), 0, s, g, static_cast<double *>(u), e);
  return hipGetLastError();
}

int norms_blocks(const DevGeom &g) {
  long long total = (long long)g.n[0] * g.n[1] * g.n[2];
  long long b = (total + 255) / 256;
  return (int)(b < kNormsMaxParts ? b : kNormsMaxParts);
}

hipError_t launch_norms(const DevGeom &g, const void *u, bool f32, double e, double *part_max,
                        long long *part_idx, double *part_sq, int *nparts, hipStream_t s) {
  int nb = norms_blocks(g);
  *nparts = nb;
  if (f32)
    hipLaunchKernelGGL(norms_kernel<float>, dim3(nb), dim3(256), 0, s, g,
                       static_cast<const float *>(u), e, part_max, part_idx, part_sq);
  else
    hipLaunchKernelGGL(norms_kernel<double>, dim3(nb), dim3(256), 0, s, g,
                       static_cast<const double *>(u), e, part_max, part_idx, part_sq);
  return hipGetLastError();
}

hipError_t launch_diagnostics(int dir, const DiagArgs &a, hipStream_t s) {
  const int lines = a.nj * a.nk;
  if (lines <= 0) return hipSuccess;
  const dim3 grid((unsigned)((lines + 255) / 256));
  if (dir == 0) hipLaunchKernelGGL(diag_kernel<0>, grid, dim3(256), 0, s, a);
  else if (dir == 1) hipLaunchKernelGGL(diag_kernel<1>, grid, dim3(256), 0, s, a);
  else if (dir == 2) hipLaunchKernelGGL(diag_kernel<2>, grid, dim3(256), 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_diag_reduce(const double *el, const double *er, const double *je, int n,
                              double *out, hipStream_t s) {
  hipLaunchKernelGGL(diag_reduce_kernel, dim3(1), dim3(1024), 0, s, el, er, je, n, out);
  return hipGetLastError();
}

hipError_t launch_tdma(int n, int nlines, const double *a, const double *b, const double *c,
                       const double *r, double *u, hipStream_t s) {
  if (n < 2 || nlines < 1) return hipErrorInvalidValue;
  int p = n / 2;
  if (p > 64) p = 64;
  const int mmax = (n + p - 1) / p;
  dim3 grid((nlines + 3) / 4);
  if (mmax <= 4) hipLaunchKernelGGL((tdma_kernel<4>), grid, dim3(256), 0, s, n, nlines, a, b, c, r, u);
  else if (mmax <= 8) hipLaunchKernelGGL((tdma_kernel<8>), grid, dim3(256), 0, s, n, nlines, a, b, c, r, u);
  else if (mmax <= 16) hipLaunchKernelGGL((tdma_kernel<16>), grid, dim3(256), 0, s, n, nlines, a, b, c, r, u);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace mts
