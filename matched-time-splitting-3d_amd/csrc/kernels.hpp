// Host-side declarations of the gfx950 kernels' launchers (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "device_types.hpp"

namespace mts {

// Device view of the grid: sizes, x pitch (row stride of the field arrays,
// a multiple of 16 doubles so every x row starts on a 128-byte line),
// node classification and the per-axis sin/cos tables.
struct DevGeom {
  int n[3];
  int px;
  const int8_t *ind;
  const double *tab[3];
};

// Device view of one direction's MIB blocks.
struct DevMib {
  const int *blk_ptr;
  const BlockHdr *hdr;
  const double *val;
  const double *cr;
};

// Right-hand-side kinds of a line stage.
enum : int {
  kRhsCopy = 0,   // d = A                                  (Douglas x; LOD-IE; TS implicit y/x)
  kRhsDouglas = 1,// d = -dt L_d S + A                      (Douglas y, z)
  kRhsCN = 2,     // d = A + theta dt (L_d A + jumps)        (LOD-CN; TS explicit x, y)
  kRhsTSZ = 3,    // CN form, then + dt/2 f(t^n) + dt/2 f(t^{n+1})   (TS z)
  kFlagLhsJump = 4,   // patched LHS rows carry theta dt vd (jumps) on their RHS
  kFlagSolve = 8,     // tridiagonal solve along the line
  kFlagEpiF = 16,     // epilogue: x + dt f(t)  on every row (LOD last sweep)
};

struct LineArgs {
  DevGeom g;
  DevMib mib;
  PhysConst P;
  const double *A;     // stage input
  const double *S;     // second input (Douglas y/z: u^n), may be null
  double *out;         // stage output (may alias A or S)
  double theta_dt;     // implicit coefficient
  double dt;
  double e_out;        // exp(-t) of the output boundary rows
  double e_in;         // exp(-t) of the input boundary rows when analytic
  int in_analytic;     // input rows 0 and n-1 replaced by the analytic boundary value
  double e_jump;       // exp(-t) of the jumps
  double e_f1, c_f1;   // TS z: f(t^n) with c = dt/2;  LOD epilogue: f(t^{n+1}) with c = dt
  double e_f2, c_f2;   // TS z: f(t^{n+1}) with c = dt/2
};

// One line stage along direction dir with the given flags.
hipError_t launch_line(int dir, int flags, const LineArgs &a, hipStream_t s);

// Douglas stage 1 right-hand side: R = dt L_y u + dt L_z u + u + dt f  (regular
// stencil everywhere; boundary nodes get the boundary value).
struct Rhs1Args {
  DevGeom g;
  PhysConst P;
  const double *u;
  double *R;
  double dt;
  double e;  // exp(-t^{n+1})
};
hipError_t launch_rhs1(const Rhs1Args &a, hipStream_t s);

// Douglas stage 1 at MIB-affected nodes: recompute R with the fictitious-point
// stencils along y and z and the jump terms of all three directions.
struct SpecArgs {
  DevGeom g;
  PhysConst P;
  DevMib mib[3];
  const int *spec;
  int nspec;
  const double *u;
  double *R;
  double dt;
  double e;
};
hipError_t launch_spec(const SpecArgs &a, hipStream_t s);

// Boundary faces: out = exact outer solution at e (+ cf * f(e) when cf != 0).
hipError_t launch_faces(const DevGeom &g, const PhysConst &P, double *out, double e, double cf,
                        hipStream_t s);

// Exact solution at e = exp(-t0) on every node (Initialization).
hipError_t launch_init(const DevGeom &g, double *u, double e, hipStream_t s);

// Error norms against the exact solution: per-block partial (max, ref index,
// sum of squares); nblocks written to *nparts.
hipError_t launch_norms(const DevGeom &g, const double *u, double e, double *part_max,
                        long long *part_idx, double *part_sq, int *nparts, hipStream_t s);
int norms_blocks(const DevGeom &g);

// Batched tridiagonal solve: nlines independent systems of length n, arrays
// line-major (line l, row i at l*n + i).  a[l*n] and c[l*n + n-1] are ignored.
hipError_t launch_tdma(int n, int nlines, const double *a, const double *b, const double *c,
                       const double *r, double *u, hipStream_t s);

}  // namespace mts
