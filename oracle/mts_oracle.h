/*
 * ORACLE (test infrastructure only): a CPU restatement in plain C of the
 * reference's per-time-step path.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, to check results against; the product
 * never does.
 *
 * It follows the reference's algorithm line by line (see mts_oracle.c for the
 * file:line of each function) and takes the interface crossings with their
 * MIB weights as input, so it checks the per-step arithmetic independently of
 * the product's kernels.  The setup that produces the crossings is pinned
 * separately against the reference build (oracle/_ref) fixtures.
 */
#ifndef MTS_ORACLE_H
#define MTS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One crossing: ints (j, k, id, left_loc), doubles coord[3], normal[3],
 * wl[9], wr[9] (unused weight entries zero). */
typedef struct {
  int n[3];
  double lo[3], hi[3];
  const signed char *ind;   /* +1 outside / -1 inside, index (iz*ny + iy)*nx + ix */
  int beta_kind;            /* 0: beta- = bp0, beta+ = bp1; 1: beta- = bp0 (1 + sin x cos y / 2) */
  double bp[2];
  int ncross[3];            /* per direction 0 = x, 1 = y, 2 = z */
  const int *cr_int[3];     /* 4 ints per crossing, sorted by line (k * nj + j) */
  const double *cr_dbl[3];  /* 24 doubles per crossing */
} mts_oracle_problem;

enum { MTS_ORACLE_ADI = 0, MTS_ORACLE_LOD_IE = 1, MTS_ORACLE_LOD_CN = 2, MTS_ORACLE_TS = 3 };

/* Exact solution at t0 (Initialization, douglas_adi.cc:1684-1714). u: nx*ny*nz, x fastest. */
int mts_oracle_init(const mts_oracle_problem *p, double t0, double *u);

/* nsteps steps of method from t0 with t += dt (helper.h:199-401); u in/out.
 * *t_end receives the final time. Returns 0, or -1 on a zero pivot / bad input. */
int mts_oracle_run(const mts_oracle_problem *p, int method, double dt, double t0, int nsteps,
                   double *u, double *t_end);

/* One Douglas-ADI step from t0 (u in/out) that also returns its stages:
 * rhs1 the x sweep's right-hand side with boundary values (douglas_adi.cc:
 * 93-359), ustar and ustar2 the outputs of the x and y sweeps (:377-514,
 * :670-807); any of them may be null. */
int mts_oracle_adi_stages(const mts_oracle_problem *p, double dt, double t0, double *u,
                          double *rhs1, double *ustar, double *ustar2);

/* Error norms (Error, douglas_adi.cc:1723-1770). */
void mts_oracle_norms(const mts_oracle_problem *p, const double *u, double t, double *linf,
                      double *l2, int argmax[3]);

/* Thomas algorithm on one line (TDMA, douglas_adi.cc:1784-1807). Returns -1 on a zero pivot. */
int mts_oracle_tdma(int n, const double *a, const double *b, const double *c, const double *r,
                    double *u);

/* Line solver used by mts_oracle_run's implicit sweeps (process-wide):
 *   THOMAS       the reference's TDMA (default),
 *   PIVOT        tridiagonal Gaussian elimination with partial pivoting
 *                (LAPACK dgtsv's algorithm): a conditioning probe,
 *   PARTITIONED  a CPU model of the GPU's line solve (segments per lane of a
 *                16/32/64-lane group, fold, cyclic reduction; Thomas for a line
 *                with a non-dominant row), operation for operation; it takes
 *                line systems, whose rows 0 and n-1 are identity rows.
 * Not part of the reference; tests use them to tell rounding from error. */
enum { MTS_ORACLE_SOLVE_THOMAS = 0, MTS_ORACLE_SOLVE_PIVOT = 1, MTS_ORACLE_SOLVE_PARTITIONED = 2 };
void mts_oracle_set_line_solver(int kind);
/* One line with the selected solver. Returns -1 on a zero pivot or bad size. */
int mts_oracle_line_solve(int n, const double *a, const double *b, const double *c,
                          const double *r, double *u);

#ifdef __cplusplus
}
#endif

#endif
