/*
 * C ABI of the MI355X Matched-TS per-time-step path.
 *
 * The reference (a single-threaded C++14 program) has no plugin ABI: its hot
 * path sits behind C++ classes.  Each entry point below replaces one of those
 * class operations; the citation names the reference interface (paths
 * relative to the reference's src/).  All functions return 0 on success and a
 * negative MTS_E_* code otherwise; mts_last_error() gives the message for the
 * calling thread.  No exception crosses this boundary.  The library owns
 * device buffers; the caller owns host buffers.  A context is not safe for
 * concurrent use from several threads.
 */
#ifndef MTS_H
#define MTS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MTS_OK = 0,
  MTS_E_ARG = -1,       /* invalid argument */
  MTS_E_GEOMETRY = -2,  /* setup refused: the reference would have exited (message says why) */
  MTS_E_DEVICE = -3,    /* HIP runtime error */
  MTS_E_STATE = -4,     /* call out of order */
  MTS_E_NOMEM = -5,
  MTS_E_DISABLED = -6   /* combination switched off in this build (message says which and why) */
};

/* temporal methods (constant.h: Temporal_Method_Type) */
enum { MTS_ADI = 0, MTS_LOD_IE = 1, MTS_LOD_CN = 2, MTS_TS = 3 };
/* field precision: FP64 (the reference's arithmetic) or the FP32 form (FP32
 * field storage and line solve, rows formed in FP64, FP64 setup and norms;
 * LOD-IE, LOD-CN and TS) */
enum { MTS_FP64 = 0, MTS_FP32 = 1 };
/* surfaces (surface/ *) */
enum { MTS_SURF_ELLIPSOID = 0, MTS_SURF_TORUS = 1, MTS_SURF_BUMPY = 2, MTS_SURF_CALLBACK = 3 };
/* diffusion coefficients: beta- = p0, beta+ = p1 (CONST); beta- = p0 (1 + sin x cos y / 2) (SINCOS) */
enum { MTS_BETA_CONST = 0, MTS_BETA_SINCOS = 1 };
/* field layouts for upload/download */
enum {
  MTS_LAYOUT_XFAST = 0, /* index (iz*ny + iy)*nx + ix (the reference's To1d, mesh.cc:93-99) */
  MTS_LAYOUT_REF = 1    /* index (ix*ny + iy)*nz + iz (the reference's CubicDoub uh[ix][iy][iz]) */
};

typedef struct {
  double lo[3], hi[3]; /* box */
  int n[3];            /* node counts per axis (the reference's x/y/z-grid) */
} mts_grid_desc;

typedef struct {
  int kind;      /* MTS_SURF_* */
  double p[8];   /* ellipsoid: a, b, c, r; torus: R, r; bumpy: R0, A, m, k */
  double (*phi)(void *user, double x, double y, double z);             /* CALLBACK */
  void (*grad)(void *user, double x, double y, double z, double g[3]); /* CALLBACK */
  void *user;
} mts_surface_desc;

typedef struct {
  int beta_kind; /* MTS_BETA_* */
  double bp[4];  /* beta parameters */
} mts_physics_desc;

/* One interface crossing, the live fields of the reference's
 * Intersection_Data (constant.h:141-156). */
typedef struct {
  int dir;          /* 0 = x, 1 = y, 2 = z line */
  int j, k;         /* the line's fixed indices: x line (iy, iz), y line (ix, iz), z line (ix, iy) */
  int id;           /* ID: 1-based count along the line, negative for a corner pair */
  int left_loc;
  int nw;           /* 6 irregular, 9 corner */
  double gamma;
  double coord[3];
  double normal[3];
  double p[9];      /* local frame rows (n, eta, tau) */
  double pinv_d0;   /* P^{-1}[dir][0] */
  double beta_in, beta_out;
  double wl[9], wr[9]; /* wei.weil[0], wei.weir[0] */
} mts_crossing;

typedef struct {
  int n[3];
  long long n_cross[3];
  int corner_pairs[3];
  int singular_warnings;
  int mib;
} mts_setup_info;

typedef struct mts_setup mts_setup;
typedef struct mts_ctx mts_ctx;
typedef struct mts_instance mts_instance;

/* Message of the last failed call on this thread ("" if none). */
const char *mts_last_error(void);
/* Library version string. */
const char *mts_version(void);

/* ---------------------------------------------------------------- setup
 * Mesh + Intersections construction on the host in FP64: node classification
 * (Mesh::Mesh mesh/mesh.cc:20-70), crossings, normals, local frames and the
 * MIB-L1/L2 fictitious-point weights (Intersections::Intersections
 * spatial/intersection.cc:22-51 and its setup chain :79-1830).
 * mib = 1 (MIB-L1, "mib-v1") or 2 (MIB-L2, "mib-v2"); accuracy = 2.
 * The work splits over host threads (MTS_SETUP_THREADS, default the hardware
 * threads up to 16; one for CALLBACK surfaces, whose callbacks are never
 * called concurrently); the result is the same on any thread count. */
int mts_setup_create(const mts_grid_desc *grid, const mts_surface_desc *surface,
                     const mts_physics_desc *physics, int mib, int accuracy, mts_setup **out);
void mts_setup_destroy(mts_setup *s);
int mts_setup_get_info(const mts_setup *s, mts_setup_info *info);
/* node classification, +1 outside / -1 inside, MTS_LAYOUT_XFAST (Mesh::mesh_value) */
int mts_setup_indicator(const mts_setup *s, int8_t *out, size_t count);
/* node coordinates along axis a (Mesh::xi/yi/zi) */
int mts_setup_coords(const mts_setup *s, int axis, double *out, size_t count);
/* crossings of direction dir in line order (Intersections::ifpx/ifpy/ifpz) */
int mts_setup_crossings(const mts_setup *s, int dir, mts_crossing *out, size_t count);

/* End-of-run interface diagnostics at time t for the manufactured solution:
 * fictitious-point errors fp = {Lmax, L2} (Refresh_Fp + Error_Fp,
 * spatial/intersection.cc:1841-2281 and 4667-4771: each fictitious point from
 * its weights, the exact solution and the exact jumps, relative to the other
 * side's exact value; two per crossing) and jump errors jump = {Lmax, L2}
 * (Refresh_Jump's jump.err and Error_Jump, :2294-2490 and 4594-4665: the GFM
 * u_dir against the exact [beta u_d], relative).  Host FP64; the relative
 * errors are infinite where the reference divides by an exact zero. */
int mts_setup_diagnostics(const mts_setup *s, double t, double fp[2], double jump[2]);
/* The same per crossing of direction dir (line order): the two fictitious
 * points' errors and the signed jump error. */
int mts_setup_crossing_errors(const mts_setup *s, int dir, double t, double *errl, double *errr,
                              double *jerr, size_t count);

/* The same diagnostics on the device, from the instance's MIB tables (its
 * setup's crossings, weights and jump records): one thread per line, then
 * one workgroup's reduction.  Per crossing bit for bit the host's
 * (mts_setup_crossing_errors); the maxima equal the host's, the L2 norms
 * agree to rounding (another summation order).  The first call allocates the
 * per-crossing buffers (3 doubles per crossing) on the instance. */
int mts_instance_diagnostics(mts_instance *inst, double t, double fp[2], double jump[2]);
int mts_instance_crossing_errors(mts_instance *inst, int dir, double t, double *errl,
                                 double *errr, double *jerr, size_t count);

/* Where the reference itself would have stopped a run that this library
 * completes.  Its set-up prepares numeric tangential flux-jump approximations
 * for every crossing (Auxiliary_eta_* / Auxiliary_tau_*,
 * spatial/intersection.cc:2746-3660) and every step evaluates them (Eta_* /
 * Tau_*, :3670-4438); their results reach no output (approximate-jump mode,
 * constant.h:16), so they are not computed here, but their geometric checks
 * call exit(0).  kinds: 0 an auxiliary point outside the domain (set-up
 * exit), 1 no three same-side nodes on both neighbouring lines (set-up exit),
 * 2 an auxiliary point on a grid node or undefined (the reference leaves its
 * node index unset: undefined behaviour, no exit), 3 a zero tangent-frame
 * component (exit at the first step).  first_kind is the kind of the first
 * exit in the reference's order (-1: it would run); first_dir / first_index
 * name that crossing (line order, as mts_setup_crossings) and message holds
 * the text the reference prints there. */
typedef struct {
  int count[4];       /* crossings per kind */
  int first_kind;
  int first_dir, first_index;
  char message[160];
} mts_reference_stops;
int mts_setup_reference_stops(const mts_setup *s, mts_reference_stops *out);

/* Node classification alone, the Mesh constructor (mesh/mesh.cc:20-70):
 * +1 outside, -1 inside, MTS_LAYOUT_XFAST; MTS_E_GEOMETRY when the interface
 * passes through a node (where the reference exits). */
int mts_mesh_classify(const mts_grid_desc *grid, const mts_surface_desc *surface, int8_t *out,
                      size_t count);

/* ------------------------------------------------------------ dense LU
 * The reference's LU class (decomposition/lu.h:7-21, lu.cc:15-196): Crout
 * decomposition with implicit-scaled partial pivoting (Numerical Recipes
 * ludcmp/lubksb), a zero pivot replaced by 1e-40.  Matrices are n*n
 * row-major; a matrix with a zero row is MTS_E_ARG ("Singular matrix in
 * LUdcmp", where the reference exits). */
typedef struct mts_lu mts_lu;
int mts_lu_create(int n, const double *a, mts_lu **out);     /* LU::LU (lu.cc:15-79) */
void mts_lu_destroy(mts_lu *lu);
int mts_lu_solve(const mts_lu *lu, const double *b, double *x); /* LU::solve(Vec) (lu.cc:82-121); b may alias x */
/* LU::solve(Matrix) (lu.cc:124-143): b and x are n*m row-major, solved column by column */
int mts_lu_solve_matrix(const mts_lu *lu, int m, const double *b, double *x);
int mts_lu_inverse(const mts_lu *lu, double *ainv);            /* LU::inverse (lu.cc:146-162) */
int mts_lu_det(const mts_lu *lu, double *det);                 /* LU::det (lu.cc:165-174) */
/* LU::mprove (lu.cc:179-196): one step of iterative refinement of x for A x = b,
 * the residual accumulated in long double */
int mts_lu_mprove(const mts_lu *lu, const double *b, double *x);

/* ---------------------------------------------------------------- device */
int mts_ctx_create(int device, mts_ctx **out);
void mts_ctx_destroy(mts_ctx *ctx);
int mts_ctx_device_info(mts_ctx *ctx, char *name, size_t name_len, size_t *hbm_bytes);

/* A device-resident solver for one problem: the constructor of Douglas_ADI
 * (temporal/douglas_adi.cc:15-53), LOD (temporal/lod.cc) or TS
 * (temporal/trapezoidal.cc).  dt is the time step; the LHS tables are built
 * for it. */
int mts_instance_create(mts_ctx *ctx, const mts_setup *setup, const mts_physics_desc *physics,
                        int method, double dt, mts_instance **out);
void mts_instance_destroy(mts_instance *inst);
/* mts_instance_create with a field precision (MTS_FP64 or MTS_FP32).  No
 * counterpart in the reference (FP64 throughout); the FP32 form halves the
 * field's HBM traffic and registers.  Every method; Douglas-ADI's stage-1
 * right-hand side is formed in FP64 from the FP32 field and rounded once. */
int mts_instance_create_ex(mts_ctx *ctx, const mts_setup *setup, const mts_physics_desc *physics,
                           int method, double dt, int precision, mts_instance **out);
int mts_instance_precision(const mts_instance *inst, int *precision);

/* Line-split autotuner.  No counterpart in the reference (its line solves are
 * sequential Thomas sweeps).  Times `steps` steps (after two untimed ones,
 * from the initial field) of this setup, method, dt and precision on a
 * scratch instance for each line split that fits the grid: lane groups of 16,
 * 32 or 64 lanes for every line (min_lanes) and 16, 32 or 64 for the general
 * tiles (gen_lanes >= min_lanes), and returns the fastest.  The choice is kept
 * in ctx for this grid, method and precision: every later instance created
 * in ctx (mts_instance_create*, the server) uses it.  Every split computes the
 * same solution to rounding.  Without a choice instances use 16 / 16, or the
 * MTS_LINE_LANES / MTS_GEN_LANES environment overrides. */
int mts_tune_lanes(mts_ctx *ctx, const mts_setup *setup, const mts_physics_desc *physics,
                   int method, double dt, int precision, int steps, int *min_lanes,
                   int *gen_lanes, double *ms_per_step);
/* The line split an instance runs (min_lanes, gen_lanes). */
int mts_instance_lanes(const mts_instance *inst, int *min_lanes, int *gen_lanes);

/* Field transfer (host <-> the instance's solution field). */
int mts_field_upload(mts_instance *inst, const double *host, int layout);
int mts_field_download(mts_instance *inst, double *host, int layout);

/* Exact solution at t0 on every node (Initialization, douglas_adi.cc:1684-1714). */
int mts_initialize(mts_instance *inst, double t0);

/* One time step ending at t_next (Solve_2nd: douglas_adi.cc:66-1110,
 * lod.cc:44-205, trapezoidal.cc:46-272).  TS evaluates its three equations at
 * t_next - dt, t_next - dt/2 and t_next as the reference's driver does
 * (helper.h:371-383).  Asynchronous on the instance's stream. */
int mts_step(mts_instance *inst, double t_next);

/* The reference's time loop (helper.h:199-238, 280-319, 361-401):
 * steps = (int)(t_finish / dt), t += dt per step, starting from t0.
 * Returns the number of steps taken in *steps and the final time in *t_end. */
int mts_run(mts_instance *inst, double t0, double t_finish, int *steps, double *t_end);

/* mts_run with the whole time loop captured once as a HIP graph and launched
 * as one unit (same kernels, same arguments, same results).  The graph is kept
 * in the instance and replayed while t0 and t_finish stay the same.  Needs
 * per-stage profiling off.  No counterpart in the reference: it removes the
 * host's per-kernel launch cost, which bounds small grids. */
int mts_run_graph(mts_instance *inst, double t0, double t_finish, int *steps, double *t_end);

/* Wait for the instance's queued work. */
int mts_sync(mts_instance *inst);

/* L-infinity (with its first argmax in the reference's loop order) and
 * L2 = sqrt(sum e^2 / (nx ny nz)) of u - exact at time t
 * (Error, douglas_adi.cc:1723-1770). */
int mts_error_norms(mts_instance *inst, double t, double *linf, double *l2, int argmax[3]);

/* mts_error_norms without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The per-block partials and their
 * reduction live in a fixed device array of the library (48 KiB per GPU, held
 * until the process ends; at most 2048 partials, which covers every grid),
 * not in the instance, and the reduction runs on the device.  Same results as
 * mts_error_norms, bit for bit.  All calls on one GPU share that array: calls
 * that overlap on a GPU (other host threads, other instances' streams) race. */
int mts_error_norms_noalloc(mts_instance *inst, double t, double *linf, double *l2,
                            int argmax[3]);

/* Device pointer of the solution field (x pitch in *pitch, in elements) and
 * the instance's HIP stream, for callers that keep data on the device.  For an
 * MTS_FP32 instance the pointer addresses floats. */
int mts_field_device(mts_instance *inst, double **dptr, int *pitch, void **stream);

/* Per-stage timing with HIP events on the instance's stream (the reference
 * only times the whole run, main.cc:119-124).  When enabled, every launch of
 * mts_step is bracketed by events; mts_profile_read synchronizes and returns
 * per-stage launch counts and total milliseconds. */
typedef struct {
  char name[32];
  int launches;
  double total_ms;
} mts_stage_time;
int mts_profile_enable(mts_instance *inst, int on);
int mts_profile_reset(mts_instance *inst);
int mts_profile_read(mts_instance *inst, mts_stage_time *out, int capacity, int *count);

/* The instance's time step. */
int mts_instance_dt(const mts_instance *inst, double *dt);
/* Device bytes the instance holds (field, MIB tables, stage records, scratch). */
int mts_instance_device_bytes(const mts_instance *inst, size_t *bytes);

/* ------------------------------------------------------------ batching
 * Advance several instances over their time loops ((int)(t_finish[i] / dt)
 * steps with t += dt from t0[i], as mts_run), issuing their steps
 * interleaved so independent solves overlap on the device (each instance
 * runs on its own stream).  No counterpart in the reference, which runs one
 * solve per process; replaces a loop of mts_run calls. */
int mts_batch_run(mts_instance *const *insts, int count, const double *t0, const double *t_finish,
                  int *steps, double *t_end);

/* mts_batch_run with one launch per line sweep across instances: the
 * instances of one method and precision step in lockstep (each with its own
 * dt and t; those with fewer steps drop out), and each solve sweep of a step
 * runs as one launch over the tiles of all of them (one per group of
 * instances whose line split and register bound match), on the first
 * instance's stream; the other launches of a step (Douglas stage 1, the
 * boundary faces, TS's explicit sweeps) follow per instance on that stream.
 * Groups of other methods or precisions run one after another.  The same
 * kernels' arithmetic as mts_run, bit for bit.  No counterpart in the
 * reference. */
int mts_batch_run_multi(mts_instance *const *insts, int count, const double *t0,
                        const double *t_finish, int *steps, double *t_end);

/* mts_batch_run with each instance's time loop as one graph (mts_run_graph),
 * the graphs launched back to back on the instances' streams. */
int mts_batch_run_graph(mts_instance *const *insts, int count, const double *t0,
                        const double *t_finish, int *steps, double *t_end);

/* ------------------------------------------------------------- serving
 * A request loop with continuous batching: submitted solves get their host
 * setup on setup threads, are admitted in arrival order while the sum of
 * their device footprints stays under hbm_cap_bytes and at most
 * max_inflight are in flight, and advance one step per round each, so
 * solves that arrive while others run join between steps.  Fields and MIB
 * tables stay on the device for a solve's lifetime.  No counterpart in the
 * reference (one solve per process, helper.h:199-401 loops). */
typedef struct {
  mts_grid_desc grid;
  mts_surface_desc surface;   /* CALLBACK surfaces are called from a setup thread */
  mts_physics_desc physics;
  int mib;                    /* 1 or 2 */
  int method;                 /* MTS_ADI .. MTS_TS */
  double dt, t0, t_finish;
  long long tag;              /* returned in the result */
  int precision;              /* MTS_FP64 or MTS_FP32 */
} mts_request;

typedef struct {
  long long id, tag;
  int status;                 /* MTS_OK or an MTS_E_* code */
  char message[160];
  int steps;
  double t_end, linf, l2;     /* error norms against the exact solution at t_end */
  int argmax[3];
  double setup_ms;            /* host setup */
  double queue_ms;            /* submit -> admitted to the device */
  double run_ms;              /* admitted -> done */
  double latency_ms;          /* submit -> done */
  size_t footprint;           /* device bytes charged against the cap */
} mts_result;

typedef struct mts_server mts_server;

int mts_server_create(mts_ctx *ctx, size_t hbm_cap_bytes, int max_inflight, int setup_threads,
                      mts_server **out);
/* Thread-safe; returns at once with the request id. */
int mts_server_submit(mts_server *s, const mts_request *req, long long *id);
/* Moves up to capacity finished results (in completion order) to out. */
int mts_server_poll(mts_server *s, mts_result *out, int capacity, int *count);
/* Blocks until at most max_outstanding submitted solves are unfinished. */
int mts_server_wait(mts_server *s, long long max_outstanding);
/* Stops the loop; solves not yet finished are dropped. */
void mts_server_destroy(mts_server *s);

/* Batched tridiagonal solve (the reference's TDMA, douglas_adi.cc:1784-1807,
 * for many lines at once): nlines systems of length n, host arrays line-major;
 * a[l*n] and c[l*n + n - 1] are ignored. */
int mts_tdma_batched(mts_ctx *ctx, int n, int nlines, const double *a, const double *b,
                     const double *c, const double *r, double *u);

#ifdef __cplusplus
}
#endif

#endif /* MTS_H */
