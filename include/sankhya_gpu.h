/* sankhya_gpu -- PDHG/PDLP engine for bounded LP and convex QP on AMD Instinct
 * MI355X (gfx950).
 *
 * Plain C ABI: every entry point is extern "C", takes plain pointers and
 * sizes, reports failure through a return code plus a thread-local message,
 * and never lets an exception or a HIP type cross the boundary.
 *
 * Each declaration cites the reference interface it replaces
 * (Jayaragul/The-solver, paths relative to that repository).
 *
 * Threading: the library queues all device work on one stream per device
 * (default priority).  The host set-up of large problems (transposes,
 * scaling sweeps, permuted copies, batch packing) runs on OpenMP threads
 * (OMP_NUM_THREADS), each output written by one thread in the serial order,
 * so results do not depend on the thread count; the registry behind the
 * CSR handles takes a lock.  A CSR handle or a PDLP session must not be
 * used from two host threads at once, and calls that run on one device
 * share its stream and event timeline: serialize them.  The caller selects
 * the device (hipSetDevice) as with the reference's cudaSetDevice
 * (apps/sankhya_pdhg.cu:79).
 */
#ifndef SANKHYA_GPU_H
#define SANKHYA_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------
 * Model / solution / options types.
 *
 * Same field order, types and enum values as c/include/sankhya.h:17-161, so a
 * caller holding an sk_model built by the reference C core (sk_read_mps,
 * sk_model_alloc) passes it here unchanged.  When sankhya.h was included
 * first its definitions are used instead of these.
 * ---------------------------------------------------------------------- */
#ifndef SANKHYA_H

typedef enum sk_status {            /* c/include/sankhya.h:17-26 */
    SK_OK = 0,
    SK_ERR_ARG = 1,
    SK_ERR_MEMORY = 2,
    SK_ERR_STRUCTURE = 3,
    SK_ERR_IO = 4,
    SK_ERR_SINGULAR = 5,
    SK_ERR_UNSUPPORTED = 6,
    SK_ERR_NUMERIC = 7
} sk_status;

typedef enum sk_result {            /* c/include/sankhya.h:29-38 */
    SK_RESULT_OPTIMAL = 0,
    SK_RESULT_INFEASIBLE = 1,
    SK_RESULT_UNBOUNDED = 2,
    SK_RESULT_ITERATION_LIMIT = 3,
    SK_RESULT_TIME_LIMIT = 4,
    SK_RESULT_NUMERIC_FAILURE = 5,
    SK_RESULT_NOT_RUN = 6,
    SK_RESULT_GAP_LIMIT = 7
} sk_result;

typedef enum sk_vartype { SK_CONTINUOUS = 0, SK_INTEGER = 1 } sk_vartype;

enum { SK_LP_AUTO = 0, SK_LP_SIMPLEX = 1, SK_LP_FIRST_ORDER = 2 };

#define SK_INFINITY 1.0e30
#define SK_IS_INF(x)     ((x) >=  SK_INFINITY)
#define SK_IS_NEG_INF(x) ((x) <= -SK_INFINITY)

typedef struct sk_csc {             /* c/include/sankhya.h:64-71 */
    int     nrow;
    int     ncol;
    int    *p;
    int    *i;
    double *x;
    int     nzmax;
} sk_csc;

typedef struct sk_model {           /* c/include/sankhya.h:82-98 */
    int         nrow;
    int         ncol;
    sk_csc      A;
    sk_csc     *Q;
    double     *c;
    double     *rlow;
    double     *rupp;
    double     *clow;
    double     *cupp;
    unsigned char *vartype;
    double      objshift;
    double      objsense;
    char      **rowname;
    char      **colname;
    char       *name;
} sk_model;

typedef struct sk_solution {        /* c/include/sankhya.h:114-132 */
    sk_result result;
    double    objective;
    double   *x;
    double   *y;
    double   *rc;
    double   *rowact;
    int       ncol;
    int       nrow;
    long long iterations;
    long long nodes;
    double    solve_seconds;
    double    primal_infeasibility;
    double    dual_infeasibility;
    double    complementarity;
    double    mip_gap;
    double    dual_bound;
} sk_solution;

typedef struct sk_options {         /* c/include/sankhya.h:139-161 */
    double primal_tol;
    double dual_tol;
    double pivot_tol;
    double time_limit;
    long long iteration_limit;
    long long node_limit;
    int    presolve;
    int    scaling;
    int    refactor_interval;
    int    verbosity;
    int    threads;
    double mip_gap_abs;
    double mip_gap_rel;
    int    mip_cuts;
    int    mip_heuristics;
    int    dual_simplex;
    int    lp_engine;
    unsigned long long random_seed;
} sk_options;

#endif /* SANKHYA_H */

/* Model/solution helpers with the reference's semantics.  Exported under
 * their own names so that this library never clashes with a linked
 * reference C core. */
void      sk_gpu_model_init(sk_model *m);                       /* sk_core.c:74-78   sk_model_init     */
void      sk_gpu_model_free(sk_model *m);                       /* sk_core.c:88-101  sk_model_free     */
sk_status sk_gpu_model_alloc(sk_model *m, int nrow, int ncol, int nzmax); /* sk_core.c:103-124 */
sk_status sk_gpu_model_validate(const sk_model *m);             /* sk_core.c:140-155 sk_model_validate */
void      sk_gpu_solution_init(sk_solution *s);                 /* sk_core.c:166-173 sk_solution_init  */
void      sk_gpu_solution_free(sk_solution *s);                 /* sk_core.c:175-180 sk_solution_free  */
void      sk_gpu_options_default(sk_options *o);                /* sk_core.c:184-207 sk_options_default*/
const char *sk_gpu_result_name(sk_result r);                    /* sk_core.c:31-43   sk_result_name    */
const char *sk_gpu_status_name(sk_status s);                    /* sk_core.c:45-58   sk_status_name    */
/* MPS / QPS reader (c/src/sk_core.c:318-378 sk_read_mps over
 * c/src/sankhya_mps.c:150-331): free format with a fixed-format fallback;
 * ROWS N/L/G/E (later N rows are free rows), COLUMNS with MARKER
 * INTORG/INTEND, RHS (objective RHS -> objshift = -value), RANGES, BOUNDS
 * LO/UP/FX/FR/MI/PL/BV/LI/UI, OBJSENSE MAX (normalised to minimisation),
 * QUADOBJ/QMATRIX/QSECTION into the full symmetric Q.  Infinite or
 * |bound| >= 1e29 becomes +-SK_INFINITY.  The model owns malloc'd arrays
 * and names: release with sk_gpu_model_free.  SK_ERR_IO when the file
 * cannot be read as MPS, SK_ERR_STRUCTURE for a bad quadratic section. */
sk_status sk_gpu_read_mps(const char *path, sk_model *m);

/* ------------------------------------------------------------------------
 * Device-resident sparse matrix.
 * Replaces SankhyaCudaCSR (cuda/include/sankhya_cuda.h:14-21) with the same
 * six fields in the same order (a SankhyaCudaCSR can be passed as is);
 * `values` are FP64.  Behind the handle the library also keeps CSR(A^T) on
 * the device, so that A^T*y is a gather, not an atomic scatter, and the
 * row-tile plan of the SpMV kernels; it finds them through
 * device_row_offsets, which identifies the handle.
 * ---------------------------------------------------------------------- */
typedef struct sk_gpu_csr {
    int   rows;
    int   cols;
    int   nnz;
    void *device_row_offsets;       /* int32[rows+1]                      */
    void *device_column_indices;    /* int32[nnz]                         */
    void *device_values;            /* double[nnz]                        */
} sk_gpu_csr;

/* cuda/include/sankhya_cuda.h:25-32 sankhya_cuda_csr_create.  Validates
 * (canonical offsets, in-range columns, finite values) before any device
 * allocation; frees previous content this library created (other field
 * values are just overwritten); 0 on success, -1 on failure. */
int  sk_gpu_csr_create(sk_gpu_csr *matrix, int rows, int cols, int nnz,
                       const int *row_offsets, const int *column_indices,
                       const double *values);
/* cuda/include/sankhya_cuda.h:33 sankhya_cuda_csr_destroy.  Idempotent;
 * a handle the library did not create (or already destroyed) is only
 * cleared. */
void sk_gpu_csr_destroy(sk_gpu_csr *matrix);

/* cuda/include/sankhya_cuda.h:38-42.  Device pointers; enqueued on the
 * library stream, return after launch.  That stream is a blocking stream,
 * so, as with the reference's default-stream contract (sankhya_cuda.h:
 * 35-37), work the caller queued on the null stream before the call (e.g.
 * a hipMemcpyAsync into x) runs first, and a later null-stream hipMemcpy of
 * y waits for the product; no sk_gpu_synchronize is needed for that.
 * 0, -1 (bad argument) or a positive hipError_t.  The transpose product is
 * deterministic (no atomics). */
int sk_gpu_spmv_device_f64(const sk_gpu_csr *matrix, const double *device_x,
                           double *device_y);
int sk_gpu_spmv_transpose_device_f64(const sk_gpu_csr *matrix,
                                     const double *device_y, double *device_x);
int sk_gpu_axpy_device_f64(int n, double alpha, const double *device_x,
                           double *device_y);
/* Waits for all work the library has queued on the current device. */
int sk_gpu_synchronize(void);

/* cuda/include/sankhya_cuda.h:153-164: host-memory convenience calls. */
int sk_gpu_spmv_f64(int rows, int cols, int nnz, const int *row_offsets,
                    const int *column_indices, const double *values,
                    const double *x, double *y);
int sk_gpu_axpy_f64(int n, double alpha, const double *x, double *y);
/* FP32 counterpart of sankhya_cuda_spmv_f64 (sankhya_cuda.h:153-158; no FP32
 * product there): the mixed engine's iteration product on host arrays --
 * values rounded to FP32 once, FP32 x, FP32 products and sums.
 * transpose = 1 computes y = A^T x through the stored CSR(A^T), as the
 * engine's A^T y does (x has `rows` entries, y `cols`).  0 or -1. */
int sk_gpu_spmv_f32(int rows, int cols, int nnz, const int *row_offsets,
                    const int *column_indices, const double *values, int transpose,
                    const float *x, float *y);

/* cuda/include/sankhya_cuda.h:167: thread-local text of the last failure. */
const char *sk_gpu_last_error(void);

/* Device diagnostic, no reference counterpart: the random-gather issue
 * rate of the current device -- `count` (at most 2^30) 4-byte gathers at
 * uniformly random positions of a table of `table_elems` floats, streamed
 * like a CSR product's (value, index) pairs, eight gathers in flight per
 * lane.  It bounds every sparse product here, whose gathered vectors are
 * indexed at random.  *gathers_per_second is the mean over `reps` timed
 * launches.  0 or -1. */
int sk_gpu_gather_rate(int table_elems, int count, int reps, double *gathers_per_second);

/* ------------------------------------------------------------------------
 * Plain PDHG with host-visible checkpoints (the reference's sankhya_cuda_*
 * solver family, cuda/src/sankhya_cuda_lp.cu:101-437).
 * ---------------------------------------------------------------------- */
typedef struct sk_gpu_lp_settings {  /* cuda/include/sankhya_cuda.h:44-52 */
    int    max_iterations;
    int    check_every;
    double tau;
    double sigma;
    double theta;
    double tolerance;
    double time_limit;               /* seconds, <= 0 disables           */
} sk_gpu_lp_settings;

typedef struct sk_gpu_lp_result {    /* cuda/include/sankhya_cuda.h:54-68 */
    int    status;                   /* 0 converged, 1 iteration limit, 2 time limit, -2 numeric */
    int    iterations;
    double objective;
    double maximum_row_violation;
    double maximum_step;
    double maximum_kkt_residual;
    double maximum_dual_residual;
    double maximum_complementarity;
} sk_gpu_lp_result;

/* cuda/include/sankhya_cuda.h:72-81 sankhya_cuda_lp_pdhg */
int sk_gpu_lp_pdhg(const sk_gpu_csr *matrix, const double *c,
                   const double *row_lower, const double *row_upper,
                   const double *col_lower, const double *col_upper,
                   sk_gpu_lp_settings settings, double *solution,
                   sk_gpu_lp_result *result);
/* cuda/include/sankhya_cuda.h:85-99 sankhya_cuda_qp_pdhg_preconditioned_with_dual */
int sk_gpu_qp_pdhg_preconditioned_with_dual(
    const sk_gpu_csr *matrix, const sk_gpu_csr *hessian,
    const double *quadratic_diagonal, const double *c,
    const double *row_lower, const double *row_upper,
    const double *col_lower, const double *col_upper,
    const double *primal_steps, const double *dual_steps,
    sk_gpu_lp_settings settings, double *solution, double *dual_solution,
    sk_gpu_lp_result *result);
/* cuda/include/sankhya_cuda.h:106-116 sankhya_cuda_diagonal_qp_pdhg */
int sk_gpu_diagonal_qp_pdhg(const sk_gpu_csr *matrix,
                            const double *quadratic_diagonal, const double *c,
                            const double *row_lower, const double *row_upper,
                            const double *col_lower, const double *col_upper,
                            sk_gpu_lp_settings settings, double *solution,
                            sk_gpu_lp_result *result);
/* cuda/include/sankhya_cuda.h:121-131 sankhya_cuda_sparse_qp_pdhg */
int sk_gpu_sparse_qp_pdhg(const sk_gpu_csr *matrix, const sk_gpu_csr *hessian,
                          const double *c, const double *row_lower,
                          const double *row_upper, const double *col_lower,
                          const double *col_upper, sk_gpu_lp_settings settings,
                          double *solution, sk_gpu_lp_result *result);
/* cuda/include/sankhya_cuda.h:137-150 sankhya_cuda_qp_pdhg_preconditioned */
int sk_gpu_qp_pdhg_preconditioned(
    const sk_gpu_csr *matrix, const sk_gpu_csr *hessian,
    const double *quadratic_diagonal, const double *c,
    const double *row_lower, const double *row_upper,
    const double *col_lower, const double *col_upper,
    const double *primal_steps, const double *dual_steps,
    sk_gpu_lp_settings settings, double *solution, sk_gpu_lp_result *result);

/* ------------------------------------------------------------------------
 * Restarted-average PDHG (PDLP).  Replaces sihps::GpuPdlp
 * (src/cuda/GpuPdlp.hpp:19-196) together with its scaling front end
 * run_first_order (src/lp/LpSolver.cpp:62-99).
 * ---------------------------------------------------------------------- */
enum {
    SK_GPU_PREC_FP64  = 0,   /* FP64 iterates and SpMV ("plain FP64")      */
    SK_GPU_PREC_MIXED = 1    /* FP32 iterates and SpMV, FP64 checks + tail */
};

typedef struct sk_gpu_pdlp_params {  /* src/cuda/GpuPdlp.hpp:19-67 PdlpParams */
    double eps_optimal;      /* relative KKT tolerance (default 1e-6)       */
    int    max_iterations;   /* default 500000                              */
    int    restart_period;   /* iterations between host checks (64)         */
    double time_limit_seconds;
    int    power_iterations; /* ||A||_2 power iterations (30)               */
    int    adaptive_step;    /* Applegate et al. 2021 s3.1 (1)              */
    double eta_min_factor;   /* 1e-6                                        */
    double eta_max_factor;   /* 1e4                                         */
    int    precision;        /* SK_GPU_PREC_*                               */
    double tail_switch;      /* MIXED, refine 0: KKT at which FP64 takes
                                over (1e-5)                                 */
    int    tail_stall_windows; /* MIXED, refine 0: windows without progress
                                  before the FP64 tail starts early (32)    */
    int    scaling;          /* 1: Ruiz then Pock-Chambolle, 0: none        */
    int    ruiz_iterations;  /* 100                                         */
    double ruiz_tol;         /* 1e-2                                        */
    int    use_graph;        /* capture each restart window in a hipGraph   */
    int    verbosity;        /* stderr: 1 set-up and autotuner timings; 2 also
                                every window's merits, restart (criterion,
                                target point, merit) and weight update     */
    double primal_weight;    /* initial primal weight; 0: ||c||/||b|| of the
                                scaled problem (no reference counterpart:
                                warm starts pass the family's learned value) */
    int    unfused;          /* 1: never fuse the step decision into the next
                                dual prox (the result is identical either way;
                                fusion saves a launch per iteration when the
                                products have few workgroups)             */
    int    refine;           /* MIXED: 1 (default) anchored refinement -- the
                                FP32 iterates hold the correction to an FP64
                                anchor that absorbs the restart point at
                                every restart, so FP32 rounding scales with
                                the distance still to go, and no FP64 tail
                                runs.  0: FP32 iterates of the full point,
                                FP64 tail from tail_switch / a stall.       */
    const volatile int *cancel; /* GpuPdlp.hpp:60-66 cancel: polled at every
                                window read-back; nonzero stops the solve
                                with stats->cancelled = 1.  NULL: never.    */
    int    restart_rule;     /* SK_GPU_RESTART_KKT (0, the reference's
                                relative-KKT merit, GpuPdlp.cpp:437-451) or
                                SK_GPU_RESTART_NDG (1, normalized duality
                                gap evaluated in FP64; LPs only).  On LPs
                                where KKT restarts stall short of the
                                tolerance (the 1M x 2M power-law LP: a
                                relative gap of 1e-3 after 900 s) NDG
                                restarts can converge (1e-6 in 310 s)      */
    int    autotune;         /* 1: at session creation, time a short run of
                                the adaptive loop for each candidate of the
                                products' row-tile entry cap {256 .. 2048},
                                grid cap {1024, 4096} and streaming matrix
                                loads {off, on}, then for an LP the best of
                                them with an LDS copy of x's longest
                                columns {32, 64 KB} in the step probe on
                                {512, 1024} workgroups, and keep the
                                fastest.  The
                                choice changes the summation order, so
                                results are reproducible for a fixed
                                choice, not across tunings.  0 (default):
                                the fixed rule (largest cap giving 1,024
                                tiles, 4,096 workgroups, streaming above
                                64 MB of FP64 entries), unless the three
                                fields below fix a tiling.                  */
    int    tile_nnz;         /* autotune 0: row-tile entry cap, 256..2048
                                (0: the rule)                               */
    int    spmv_grid;        /* autotune 0: workgroup cap of the products,
                                1..4096 (0: 4096)                           */
    int    stream_loads;     /* autotune 0: 0 the rule, 1 off, 2 on         */
    int    probe_grid;       /* autotune 0: workgroup cap of the step probe
                                (A x, with Q x for a QP) alone, 1..4096
                                (0: spmv_grid)                              */
    int    xcache_bytes;     /* autotune 0, LP: the step probe gathers the
                                last columns of x in slot order (the
                                longest) from a copy of this many bytes in
                                each workgroup's LDS; a multiple of 16 up
                                to 65,536 (0: none)                         */
    int    pad;
} sk_gpu_pdlp_params;

enum {
    SK_GPU_RESTART_KKT = 0,
    SK_GPU_RESTART_NDG = 1
};

typedef struct sk_gpu_pdlp_stats {   /* src/cuda/GpuPdlp.hpp:69-92 PdlpStats */
    int    converged;
    int    cancelled;
    int    iterations;
    int    restarts;
    double primal_objective;
    double dual_objective;
    double relative_primal_residual;
    double relative_dual_residual;
    double relative_gap;
    double matrix_norm;
    double seconds;
    int    rejected_steps;
    double final_step_size;
    double step_size_ratio;
    int    host_syncs;
    /* additions */
    int    fp32_iterations;      /* iterations run with FP32 iterates       */
    int    fp64_iterations;      /* iterations run with FP64 iterates       */
    long long kernel_launches;
    double loop_seconds;         /* device-resident loop only               */
    double setup_seconds;        /* scaling, binning, upload                */
    double primal_weight;
    int    reanchors;            /* anchored refinement: anchor moves       */
    int    iterations_to_merit[4]; /* iterations at the first window whose
                                    relative KKT merit max(primal, dual, gap)
                                    is <= 1e-3, 1e-4, 1e-5, 1e-6 (0: never) */
    double seconds_to_merit[4];  /* loop seconds at those windows           */
    double restart_merit[2];     /* last window's restart merits of the
                                    current and the averaged point: their
                                    relative KKT (restart_rule KKT) or their
                                    normalized duality gap (NDG)            */
    int    tile_nnz;             /* row-tile entry cap of A and A^T         */
    int    spmv_grid;            /* workgroup cap of the products           */
    int    stream_loads;         /* non-temporal matrix loads (A and A^T)   */
    int    tuned;                /* the autotuner chose the tiling          */
    double tune_seconds;         /* time the autotuner took                 */
    int    probe_grid;           /* workgroup cap of the step probe         */
    int    xcache_bytes;         /* LDS copy of x in the step probe (bytes) */
} sk_gpu_pdlp_stats;

void sk_gpu_pdlp_params_default(sk_gpu_pdlp_params *p);

/* Solve min c'x (+ 1/2 x'Qx) s.t. row_lower <= A x <= row_upper,
 * col_lower <= x <= col_upper.  A is CSR in host memory (FP64); Q, if given,
 * is a full symmetric positive semidefinite CSR (n x n) and needs the
 * adaptive step.  Bounds use IEEE +-inf or the +-1e30
 * sentinel.  x0/y0 are an optional warm start (either may be NULL).  Outputs
 * x[n], y[m] in host memory, in the original (unscaled) space, with the
 * sign convention of sk_verify: rc = c + Qx + A'y.  Returns 0 (read
 * stats->converged), -1 invalid input / device failure, -2 non-finite
 * iterate. */
int sk_gpu_pdlp_solve(int rows, int cols, int nnz, const int *row_offsets,
                      const int *column_indices, const double *values,
                      const int *q_row_offsets, const int *q_column_indices,
                      const double *q_values, int q_nnz,
                      const double *c, const double *row_lower,
                      const double *row_upper, const double *col_lower,
                      const double *col_upper, const double *x0,
                      const double *y0, const sk_gpu_pdlp_params *params,
                      double *x, double *y, sk_gpu_pdlp_stats *stats);

/* Resident problem: the scaled, permuted, binned operator and its ||A||_2
 * stay on the device across solves (the role of a GpuPdlp object,
 * src/cuda/GpuPdlp.hpp:128-196).  `update` replaces c and the bounds (same
 * A) -- the parametric-family case; `iterate` runs raw iterations from a
 * cold start with no checks (benchmarking; with kernel_ms[5] it also
 * returns the mean event-timed milliseconds of the parts of an adaptive
 * iteration -- [0] dual prox, [1] A^T y SpMV + primal prox, [2] (Q x and)
 * A x SpMV + step probe, [3] accept/reject decision -- and [4] of the whole
 * iteration). */
typedef struct sk_gpu_pdlp_session sk_gpu_pdlp_session;
int  sk_gpu_pdlp_session_create(int rows, int cols, int nnz, const int *row_offsets,
                                const int *column_indices, const double *values,
                                const double *c, const double *row_lower,
                                const double *row_upper, const double *col_lower,
                                const double *col_upper, const sk_gpu_pdlp_params *params,
                                sk_gpu_pdlp_session **session);
/* Convex QP: Q is the full symmetric Hessian as CSR (n x n), PSD. */
int  sk_gpu_pdlp_session_create_qp(int rows, int cols, int nnz, const int *row_offsets,
                                   const int *column_indices, const double *values,
                                   const int *q_row_offsets, const int *q_column_indices,
                                   const double *q_values, int q_nnz, const double *c,
                                   const double *row_lower, const double *row_upper,
                                   const double *col_lower, const double *col_upper,
                                   const sk_gpu_pdlp_params *params,
                                   sk_gpu_pdlp_session **session);
int  sk_gpu_pdlp_session_update(sk_gpu_pdlp_session *session, const double *c,
                                const double *row_lower, const double *row_upper,
                                const double *col_lower, const double *col_upper);
int  sk_gpu_pdlp_session_solve(sk_gpu_pdlp_session *session, const double *x0,
                               const double *y0, double *x, double *y,
                               sk_gpu_pdlp_stats *stats);
int  sk_gpu_pdlp_session_iterate(sk_gpu_pdlp_session *session, int iterations,
                                 double *kernel_ms);
/* Continue the session's last solve from exactly where it stopped (same
 * iterate, step size, primal weight, running averages and restart state)
 * under the current limits: at most max_iterations more iterations, the
 * time limit counted from this call.  The typical use is a tighter
 * eps_optimal after a converged solve (set_limits, then resume).  Stats
 * are cumulative over the solve and its resumes.  No reference counterpart
 * (GpuPdlp::solve always starts over); -1 when no solve ran. */
int  sk_gpu_pdlp_session_resume(sk_gpu_pdlp_session *session, double *x, double *y,
                                sk_gpu_pdlp_stats *stats);
/* A warm start with the cold start as its fallback (no reference
 * counterpart: the reference always starts at zero, c/src/sk_solve.c:871).
 * Solves from (x0, y0) -- a learned prediction, say -- for at most
 * warm_iterations iterations; if that does not reach eps_optimal, solves
 * again from zero under the session's own limits (time counted from the
 * call).  *used_warm tells which start produced x, y; stats->iterations
 * counts both attempts.  The budget should carry slack over a cold solve
 * (twice the family's typical cold iteration count, say): a warm start
 * that is slightly slower than a cold one then still finishes, while one
 * that cannot converge costs a bounded detour. */
int  sk_gpu_pdlp_session_solve_fallback(sk_gpu_pdlp_session *session, const double *x0,
                                        const double *y0, int warm_iterations, double *x,
                                        double *y, sk_gpu_pdlp_stats *stats, int *used_warm);
/* The stopping limits of later solves and resumes (GpuPdlp.hpp:19-37
 * max_iterations, time_limit_seconds, eps_optimal); the resident operator
 * and the loop state are kept. */
int  sk_gpu_pdlp_session_set_limits(sk_gpu_pdlp_session *session, int max_iterations,
                                    double time_limit_seconds, double eps_optimal);
void sk_gpu_pdlp_session_destroy(sk_gpu_pdlp_session *session);

/* ------------------------------------------------------------------------
 * Batched PDLP: many small LPs in one persistent launch.  No reference
 * counterpart (the reference solves one model per call, sk_solve,
 * c/src/sk_solve.c:1009); each instance follows the same algorithm as
 * sk_gpu_pdlp_solve: with pdlp.precision SK_GPU_PREC_MIXED, FP32 iterates
 * until the relative KKT merit reaches pdlp.tail_switch (or stalls near
 * it), then a second persistent launch continues the handed-off instances
 * with FP64 iterates; with SK_GPU_PREC_FP64, FP64 iterates throughout.
 * KKT evaluation is FP64 in both.
 * ---------------------------------------------------------------------- */
typedef struct sk_gpu_batch_options {
    sk_gpu_pdlp_params pdlp;  /* eps, iteration limit, restart period, ... */
    int block_threads;        /* threads per instance workgroup; 0 = autotune */
    int use_graph;            /* replay the launch from a captured hipGraph  */
} sk_gpu_batch_options;

typedef struct sk_gpu_batch_result {
    int    status;            /* 0 relative KKT <= eps, 1 iteration limit  */
    int    iterations;        /* total, FP32 phase and FP64 tail           */
    int    restarts;
    int    fp32_iterations;   /* iterations run with FP32 iterates         */
    double relative_primal_residual;
    double relative_dual_residual;
    double relative_gap;
    double objective;         /* original space, including objshift        */
} sk_gpu_batch_result;

typedef struct sk_gpu_batch_stats {
    double setup_seconds;     /* scaling, packing, upload                  */
    double device_seconds;    /* the persistent launches (event-timed)     */
    double seconds;
    int    block_threads, grid, lds_bytes, converged;
    long long total_iterations;
    double tune_ms[3];        /* autotuner timings for 128/256/512 threads */
    long long fp32_iterations;
    int    handoffs;          /* instances finished by the FP64 tail        */
    int    grid_fp64, lds_bytes_fp64;   /* launch shape of the FP64 phase   */
    int    pad;
} sk_gpu_batch_stats;

void sk_gpu_batch_options_default(sk_gpu_batch_options *o);
/* models[count]: continuous LPs without Q.  x_out holds the instances'
 * columns back to back (sum of ncol), y_out their rows (sum of nrow).
 * An instance's vectors live in one workgroup's LDS: the largest instance
 * must fit (4 * (7 ncol + 9 nrow) + 4 * (ncol + nrow + 2) bytes for the
 * FP32 phase, within the device's per-workgroup LDS; about 160 KB on
 * MI355X), else -1 with the size in sk_gpu_last_error(); the FP64 tail is
 * skipped when its 8-byte vectors do not fit.  block_threads = 0 picks the
 * size by timing a sample (tune_ms); the result of an instance depends on
 * block_threads (the reductions' shape), not on the batch around it. */
int  sk_gpu_batch_solve(int count, const sk_model *models, const sk_gpu_batch_options *opt,
                        double *x_out, double *y_out, sk_gpu_batch_result *results,
                        sk_gpu_batch_stats *stats);

/* ------------------------------------------------------------------------
 * Learned warm start.  No reference counterpart (the reference cold-starts
 * every PDHG solve, c/src/sk_solve.c:871).  A 3-layer MLP
 * inputs -> hidden -> hidden -> outputs with ReLU, trained and evaluated on
 * the GPU by hand-written BF16 MFMA kernels with FP32 accumulation and FP32
 * master weights (Adam).  Inputs and targets are standardised with the
 * means / standard deviations given to sk_gpu_mlp_set_scaling; the loss is
 * the mean squared error of the standardised targets.  Parameter order in
 * get/set/gradients: W1 [in x hid], b1, W2 [hid x hid], b2, W3 [hid x out],
 * b3 (row-major, x @ W).
 * ---------------------------------------------------------------------- */
typedef struct sk_gpu_mlp sk_gpu_mlp;
int       sk_gpu_mlp_create(int inputs, int hidden, int outputs, unsigned long long seed,
                            sk_gpu_mlp **mlp);
void      sk_gpu_mlp_destroy(sk_gpu_mlp *mlp);
long long sk_gpu_mlp_param_count(const sk_gpu_mlp *mlp);
int       sk_gpu_mlp_set_scaling(sk_gpu_mlp *mlp, const float *in_mean, const float *in_std,
                                 const float *out_mean, const float *out_std);
int       sk_gpu_mlp_get_params(sk_gpu_mlp *mlp, float *params);
int       sk_gpu_mlp_set_params(sk_gpu_mlp *mlp, const float *params);
/* X [count x inputs], Y [count x outputs] host FP32; one Adam step per
 * batch of `batch` rows, epoch_loss[epochs] receives the mean batch loss. */
int       sk_gpu_mlp_train(sk_gpu_mlp *mlp, const float *X, const float *Y, int count, int epochs,
                           int batch, float lr, unsigned long long seed, float *epoch_loss);
int       sk_gpu_mlp_predict(sk_gpu_mlp *mlp, const float *X, int count, float *Y);
/* Loss and the flat gradient of one batch (for checking against an oracle). */
int       sk_gpu_mlp_gradients(sk_gpu_mlp *mlp, const float *X, const float *Y, int count,
                               float *loss, float *grad);

/* ------------------------------------------------------------------------
 * sk_solve-style entry point.  Mirrors sk_solve (c/include/sankhya.h:173,
 * c/src/sk_solve.c:1009-1046) for continuous LP / convex QP, running the
 * GPU engine.  OPTIMAL is reported only when, on the device in FP64, the
 * relative KKT triple of the original model is within pdlp.eps_optimal AND
 * the sk_verify certificate is within the reference's margins
 * (c/src/sk_solve.c:987-996): primal infeasibility <= 100 primal_tol, and
 * for a QP dual infeasibility and complementarity <= 100 dual_tol.  A
 * converged loop that misses either continues from its point with a 10x
 * tighter tolerance (at most three times) within the iteration and time
 * limits left; otherwise ITERATION_LIMIT or TIME_LIMIT (sk_solve.c:997-999),
 * NUMERIC_FAILURE for a non-finite point.  `g` is read only (NULL:
 * defaults); statistics go to `stats` (may be NULL).
 * ---------------------------------------------------------------------- */
typedef struct sk_gpu_options {
    sk_gpu_pdlp_params pdlp;
    const double *x0;        /* optional warm start, original space          */
    const double *y0;
} sk_gpu_options;

void      sk_gpu_solver_options_default(sk_gpu_options *g);
sk_status sk_gpu_solve(const sk_model *m, const sk_options *o,
                       const sk_gpu_options *g, sk_solution *s,
                       sk_gpu_pdlp_stats *stats);
/* c/include/sankhya.h:177 sk_verify, evaluated on the device in FP64. */
sk_status sk_gpu_verify(const sk_model *m, sk_solution *s);

/* PDLP relative KKT triple in FP64 on the device
 * (src/cuda/GpuPdlp.cpp:171-176, 226-249): out[0] primal, out[1] dual,
 * out[2] gap, out[3] primal objective, out[4] dual objective. */
int sk_gpu_relative_kkt(int rows, int cols, int nnz, const int *row_offsets,
                        const int *column_indices, const double *values,
                        const double *c, const double *row_lower,
                        const double *row_upper, const double *col_lower,
                        const double *col_upper, const double *x,
                        const double *y, double out[5]);
/* The relative KKT triple of sk_gpu_relative_kkt for an LP already on the
 * device, allocating nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  matrix:
 * a handle from sk_gpu_csr_create; c, bounds (IEEE inf or the 1e30
 * sentinel), x, y: device vectors in the original space; device_out:
 * 5 device doubles (primal, dual, gap, primal objective, dual objective).
 * Queued on the library stream; read device_out after a null-stream copy
 * or sk_gpu_synchronize.  Scratch: block partials in a static device
 * array of 256 KB per GPU (two products of at most 4,096 workgroups,
 * 4 sums each), so there is no limit on rows, columns or nonzeros; calls
 * that overlap on one GPU share that array and race.  0, -1 (bad
 * argument) or a positive hipError_t. */
int sk_gpu_relative_kkt_device_noalloc(const sk_gpu_csr *matrix, const double *device_c,
                                       const double *device_row_lower,
                                       const double *device_row_upper,
                                       const double *device_col_lower,
                                       const double *device_col_upper,
                                       const double *device_x, const double *device_y,
                                       double *device_out);

/* The same triple with an optional full symmetric Hessian (CSR, n x n):
 * primal objective c'x + x'Qx/2, reduced cost c + Qx + A'y. */
int sk_gpu_relative_kkt_qp(int rows, int cols, int nnz, const int *row_offsets,
                           const int *column_indices, const double *values,
                           const int *q_row_offsets, const int *q_column_indices,
                           const double *q_values, int q_nnz, const double *c,
                           const double *row_lower, const double *row_upper,
                           const double *col_lower, const double *col_upper, const double *x,
                           const double *y, double out[5]);

/* ------------------------------------------------------------------------
 * LP pipeline of the C++ solver: presolve -> first-order GPU solve of the
 * reduced model -> postsolve -> original-space primal gate
 * (src/lp/LpSolver.cpp:127-333 solve_lp with LpMethod::FIRST_ORDER).
 *
 * The problem is in the form of src/lp/LpProblem.hpp:15-28: A x + s = rhs
 * with slack_lower <= s <= slack_upper (so rhs - slack_upper <= A x <=
 * rhs - slack_lower) and lower <= x <= upper.  Absent bounds are IEEE
 * +-infinity: presolve and the gate read a finite 1e30 as a finite bound,
 * as the reference does (the GPU engine reads |b| >= 1e30 as absent).
 * A is CSR; row_types ('L', 'G', 'E', 'N') is diagnostic and may be NULL
 * on input.  Presolve and postsolve run on the host (their passes are
 * sequential by definition: each reduction sees the bounds the previous
 * one left); the reduced model goes to the GPU engine.
 * ---------------------------------------------------------------------- */
typedef struct sk_gpu_lp_problem {  /* src/lp/LpProblem.hpp:15-28 LpProblem */
    int     rows, cols, nnz;
    int    *row_ptr;                 /* [rows + 1]                           */
    int    *col_idx;                 /* [nnz]                                */
    double *values;                  /* [nnz]                                */
    double *obj;                     /* [cols]                               */
    double *rhs;                     /* [rows]                               */
    char   *row_types;               /* [rows] or NULL                       */
    double *lower, *upper;           /* [cols]                               */
    double *slack_lower, *slack_upper; /* [rows]                             */
} sk_gpu_lp_problem;

enum { SK_GPU_PRESOLVE_OK = 0, SK_GPU_PRESOLVE_INFEASIBLE = 1, SK_GPU_PRESOLVE_UNBOUNDED = 2 };

typedef struct sk_gpu_presolve_result {  /* src/lp/Presolve.hpp:12-55 PresolveResult */
    int    status;                   /* SK_GPU_PRESOLVE_*                    */
    const char *reason;              /* which test concluded, or NULL        */
    int    reason_row, reason_col, reason_pass;
    sk_gpu_lp_problem reduced;       /* owned by the result                  */
    int   *kept_columns;             /* reduced column k is original kept_columns[k] */
    int   *kept_rows;                /* reduced row r is original kept_rows[r] */
    double *fixed_value;             /* [original_n_cols]: value of a removed column */
    char  *column_removed;           /* [original_n_cols]                    */
    int    original_n_rows, original_n_cols;
    int    removed_rows, removed_cols;
} sk_gpu_presolve_result;

/* src/lp/Presolve.hpp:92 presolve (max_passes <= 0 means 20).  Returns 0
 * (read out->status), -1 on invalid input or allocation failure.  Release
 * with sk_gpu_presolve_free. */
int  sk_gpu_presolve(const sk_gpu_lp_problem *problem, int max_passes, sk_gpu_presolve_result *out);
/* src/lp/Presolve.hpp:95 postsolve: x[original_n_cols] from the reduced
 * solution (reduced_len entries; extra kept columns read as 0). */
int  sk_gpu_postsolve(const sk_gpu_presolve_result *result, const double *reduced_x, int reduced_len,
                      double *x);
void sk_gpu_presolve_free(sk_gpu_presolve_result *result);

/* An sk_model (continuous LP) in LpProblem form: rows with a finite upper
 * bound anchor rhs there (slack in [0, upper - lower]), rows with only a
 * lower bound anchor it at the lower bound (slack in [-inf, 0]), free rows
 * get rhs 0; the +-1e30 sentinel becomes IEEE infinity.  Owned arrays:
 * release with sk_gpu_lp_problem_free. */
int  sk_gpu_lp_problem_from_model(const sk_model *m, sk_gpu_lp_problem *p);
void sk_gpu_lp_problem_free(sk_gpu_lp_problem *p);

enum {  /* src/lp/Simplex.hpp:19 LpStatus */
    SK_GPU_LP_OPTIMAL = 0, SK_GPU_LP_INFEASIBLE = 1, SK_GPU_LP_UNBOUNDED = 2,
    SK_GPU_LP_ITERATION_LIMIT = 3, SK_GPU_LP_NUMERICAL_FAILURE = 4
};

typedef struct sk_gpu_lp_solver_options {  /* src/lp/LpSolver.hpp:46-106 (FIRST_ORDER) */
    int    use_presolve;             /* 1                                    */
    int    presolve_passes;          /* 20                                   */
    sk_gpu_pdlp_params pdlp;
} sk_gpu_lp_solver_options;

typedef struct sk_gpu_lp_solution {  /* src/lp/LpSolver.hpp:117-165 LpSolution */
    int    status;                   /* SK_GPU_LP_*                          */
    int    has_x;                    /* x[] was written (OPTIMAL or a failed gate) */
    double objective_value;          /* original costs . postsolved x        */
    int    iterations;
    double presolve_seconds;
    double solve_seconds;
    int    presolve_removed_rows;
    int    presolve_removed_cols;
    int    used_first_order;
    double primal_residual;          /* original-space gate value            */
    sk_gpu_pdlp_stats pdlp;
} sk_gpu_lp_solution;

void sk_gpu_lp_solver_options_default(sk_gpu_lp_solver_options *o);
/* src/lp/LpSolver.cpp:127 solve_lp with method FIRST_ORDER.  x[cols]
 * receives the solution in the original column space.  Returns 0 (read
 * solution->status) or -1 on invalid input / device failure. */
int  sk_gpu_solve_lp(const sk_gpu_lp_problem *problem, const sk_gpu_lp_solver_options *options,
                     double *x, sk_gpu_lp_solution *solution);
/* src/lp/LpSolver.cpp:23-50: max(row violation / (1 + ||rhs||_inf), bound
 * violation) of x against the problem, rows summed in CSR order. */
double sk_gpu_lp_primal_residual(const sk_gpu_lp_problem *problem, const double *x);

#ifdef __cplusplus
}
#endif

#endif /* SANKHYA_GPU_H */
