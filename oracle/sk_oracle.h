/* CPU oracle for the sankhya_gpu engine -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatements of the reference algorithms on the GPU PDHG path.
 * Only tests, the smoke check and the CPU-baseline leg of the benchmark load
 * this library, always as the checker; the product never calls it.
 *
 *   orc_model_validate     c/src/sk_core.c:126-155
 *   orc_verify             c/src/sk_core.c:394-510          (sk_verify)
 *   orc_interval_infeasible c/src/sk_solve.c:97-132
 *   orc_solve_first_order  c/src/sk_solve.c:801-1007, 1009-1046
 *                          (sk_solve with lp_engine = SK_LP_FIRST_ORDER)
 *   orc_cuda_pdhg          cuda/src/sankhya_cuda_lp.cu:101-385 (solve_qp)
 *   orc_ruiz               src/lp/Scaling.cpp:11-92
 *   orc_pock_chambolle     src/lp/Scaling.cpp:111-145
 *   orc_power_norm         src/cuda/GpuPdlp.cpp:179-211
 *   orc_pdlp_solve         src/cuda/GpuPdlp.cpp:252-534 + PdlpKernels.cu
 *   orc_first_order        src/lp/LpSolver.cpp:62-99 (run_first_order)
 *   orc_relative_kkt       src/cuda/GpuPdlp.cpp:171-176, 226-249 and
 *                          src/cuda/PdlpKernels.cu:101-209
 *   orc_ndg                no reference counterpart: the normalized
 *                          duality gap of Applegate et al. 2021, s.3.1,
 *                          the restart merit of restart_rule NDG
 *   orc_pdlp_solve_qp,     no reference counterpart (GpuPdlp is LP-only):
 *   orc_first_order_qp     the two above with a convex Hessian, as the
 *                          device engine extends them
 *
 * Pinning: see the header of sk_oracle.c.
 */
#ifndef SK_ORACLE_H
#define SK_ORACLE_H

#include "sankhya_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_csr {
    int rows, cols, nnz;
    int *p;        /* rows+1 */
    int *j;        /* nnz column indices */
    double *v;     /* nnz values */
} orc_csr;

int  orc_csr_from_csc(const sk_csc *a, orc_csr *out);     /* CSR(A) from CSC(A) */
int  orc_csr_transpose(const orc_csr *a, orc_csr *out);
void orc_csr_free(orc_csr *a);
void orc_csr_mv(const orc_csr *a, const double *x, double *y);

sk_status orc_model_validate(const sk_model *m);
sk_status orc_verify(const sk_model *m, sk_solution *s);
int       orc_interval_infeasible(const sk_model *m, double tolerance);
sk_status orc_solve_first_order(const sk_model *m, const sk_options *o, sk_solution *s);
void      orc_solution_free(sk_solution *s);
void      orc_options_default(sk_options *o);

int orc_cuda_pdhg(const orc_csr *A, const orc_csr *Q, const double *qdiag,
                  const double *c, const double *rl, const double *ru,
                  const double *cl, const double *cu, sk_gpu_lp_settings settings,
                  const double *primal_steps, const double *dual_steps,
                  double *x, double *y, sk_gpu_lp_result *result);

void   orc_ruiz(const orc_csr *a, int max_iters, double tol, double *R, double *C);
void   orc_pock_chambolle(const orc_csr *a, double alpha, double *R, double *C);
double orc_power_norm(const orc_csr *a, int iterations);
/* Normalized duality gap of (x, y) against the restart point (xr, yr) for
 * the LP min c'x, rl <= Ax <= ru, l <= x <= u (IEEE +-inf bounds); ax = A x,
 * g = c + A'y given; omega the primal weight.  0 when z = z_r. */
double orc_ndg(int n, int m, const double *x, const double *y, const double *ax, const double *g,
               const double *l, const double *u, const double *rl, const double *ru, const double *xr,
               const double *yr, double omega);

/* PDLP on an already scaled LP (bounds: IEEE +-inf means absent). */
int orc_pdlp_solve(const orc_csr *A, const double *c, const double *l,
                   const double *u, const double *rl, const double *ru,
                   const sk_gpu_pdlp_params *params, const double *x0,
                   const double *y0, double *x, double *y,
                   sk_gpu_pdlp_stats *stats);
/* Scaling front end + PDLP + unscaling; inputs in the original space with
 * +-1e30 or +-inf for absent bounds; outputs x, y in the original space. */
int orc_first_order(const orc_csr *A, const double *c, const double *l,
                    const double *u, const double *rl, const double *ru,
                    const sk_gpu_pdlp_params *params, double *x, double *y,
                    sk_gpu_pdlp_stats *stats);
/* The same two with a convex Hessian Q (CSR, n x n, symmetric): the
 * primal gradient is c + Qx + A'y at the current x, the step bound covers
 * |dy'A dx| + dx'Q dx / 2, the KKT terms use c + Qx and both objectives
 * carry x'Qx/2; scaling comes from A alone and Q is scaled to C Q C.  No
 * reference counterpart (GpuPdlp is LP-only); it restates the device
 * engine's QP rule.  Adaptive steps and KKT restarts only (else -1). */
int orc_pdlp_solve_qp(const orc_csr *A, const orc_csr *Q, const double *c,
                      const double *l, const double *u, const double *rl,
                      const double *ru, const sk_gpu_pdlp_params *params,
                      const double *x0, const double *y0, double *x, double *y,
                      sk_gpu_pdlp_stats *stats);
int orc_first_order_qp(const orc_csr *A, const orc_csr *Q, const double *c,
                       const double *l, const double *u, const double *rl,
                       const double *ru, const sk_gpu_pdlp_params *params,
                       double *x, double *y, sk_gpu_pdlp_stats *stats);
/* out: rel primal, rel dual, rel gap, primal objective, dual objective. */
void orc_relative_kkt(const orc_csr *A, const orc_csr *Q, const double *c,
                      const double *rl, const double *ru, const double *l,
                      const double *u, const double *x, const double *y,
                      double out[5]);

#ifdef __cplusplus
}
#endif

#endif
