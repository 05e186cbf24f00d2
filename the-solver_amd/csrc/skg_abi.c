/* Public C ABI of sankhya_gpu: device CSR handles and products, the plain
 * PDHG solver family with host-visible checkpoints, the sk_solve-style
 * entry point and the FP64 certificates.  C11; device work goes through
 * the skk_* layer. */
#define _POSIX_C_SOURCE 200809L
#include "skg_internal.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------ device CSR */

typedef struct csr_plan {
    dcsr A, AT;      /* caller row order, bins addressed through row lists */
} csr_plan;

/* The handle keeps the reference's six fields (SankhyaCudaCSR,
 * cuda/include/sankhya_cuda.h:14-21), so the plan behind it -- CSR(A^T),
 * the row tiles -- is found through its row-offset allocation, which the
 * plan owns and no other handle shares.  A few live handles at a time: a
 * locked list, searched linearly. */
typedef struct plan_slot {
    const void *key;
    csr_plan *plan;
} plan_slot;

static pthread_mutex_t g_plan_lock = PTHREAD_MUTEX_INITIALIZER;
static plan_slot *g_plans;
static int g_plan_count, g_plan_cap;

static int plan_register(csr_plan *plan)
{
    int rc = 0;
    pthread_mutex_lock(&g_plan_lock);
    if (g_plan_count == g_plan_cap) {
        const int cap = g_plan_cap ? 2 * g_plan_cap : 16;
        plan_slot *grown = (plan_slot *)realloc(g_plans, sizeof(plan_slot) * (size_t)cap);
        if (grown) { g_plans = grown; g_plan_cap = cap; }
        else rc = -1;
    }
    if (!rc) {
        g_plans[g_plan_count].key = plan->A.rp;
        g_plans[g_plan_count].plan = plan;
        ++g_plan_count;
    }
    pthread_mutex_unlock(&g_plan_lock);
    return rc;
}

/* The plan registered for `key`, removed from the list when `take`. */
static csr_plan *plan_find(const void *key, int take)
{
    csr_plan *plan = NULL;
    int k;
    if (!key) return NULL;
    pthread_mutex_lock(&g_plan_lock);
    for (k = 0; k < g_plan_count; ++k)
        if (g_plans[k].key == key) {
            plan = g_plans[k].plan;
            if (take) g_plans[k] = g_plans[--g_plan_count];
            break;
        }
    pthread_mutex_unlock(&g_plan_lock);
    return plan;
}

/* cuda/src/sankhya_cuda.cu:154-163.  Frees only what this library
 * created: a handle it does not know (zeroed, copied after a destroy, or
 * never created) is just cleared. */
void sk_gpu_csr_destroy(sk_gpu_csr *matrix)
{
    csr_plan *plan;
    if (!matrix) return;
    plan = plan_find(matrix->device_row_offsets, 1);
    if (plan) { dcsr_free(&plan->A); dcsr_free(&plan->AT); free(plan); }
    matrix->rows = matrix->cols = matrix->nnz = 0;
    matrix->device_row_offsets = matrix->device_column_indices = matrix->device_values = NULL;
}

static int plan_upload(dcsr *d, const hcsr *a)
{
    return dcsr_upload(d, a, 0);
}

/* cuda/src/sankhya_cuda.cu:116-152 sankhya_cuda_csr_create */
int sk_gpu_csr_create(sk_gpu_csr *matrix, int rows, int cols, int nnz, const int *row_offsets,
                      const int *column_indices, const double *values)
{
    hcsr a, at;
    csr_plan *plan;
    skg_clear_error();
    if (!matrix) return skg_fail("invalid persistent CSR input");
    if (hcsr_validate(rows, cols, nnz, row_offsets, column_indices, values)) return -1;
    if (skg_have_device()) return -1;
    sk_gpu_csr_destroy(matrix);
    a.rows = rows; a.cols = cols; a.nnz = nnz;
    a.p = (int *)row_offsets; a.j = (int *)column_indices; a.v = (double *)values;
    if (hcsr_transpose(&a, &at)) return skg_fail("out of host memory");
    plan = (csr_plan *)calloc(1, sizeof(csr_plan));
    if (!plan) { hcsr_free(&at); return skg_fail("out of host memory"); }
    if (plan_upload(&plan->A, &a) || plan_upload(&plan->AT, &at)) {
        hcsr_free(&at); dcsr_free(&plan->A); dcsr_free(&plan->AT); free(plan);
        return -1;
    }
    hcsr_free(&at);
    if (plan_register(plan)) {
        dcsr_free(&plan->A); dcsr_free(&plan->AT); free(plan);
        return skg_fail("out of host memory");
    }
    matrix->rows = rows; matrix->cols = cols; matrix->nnz = nnz;
    matrix->device_row_offsets = plan->A.rp;
    matrix->device_column_indices = plan->A.ci;
    matrix->device_values = plan->A.vd;
    return 0;
}

static const csr_plan *plan_of(const sk_gpu_csr *m)
{
    return m ? plan_find(m->device_row_offsets, 0) : NULL;
}

int sk_gpu_spmv_device_f64(const sk_gpu_csr *matrix, const double *dx, double *dy)
{                                                  /* cuda/src/sankhya_cuda.cu:165-170 */
    const csr_plan *p = plan_of(matrix);
    int rc;
    skg_clear_error();
    if (!matrix) return skg_fail("null device CSR pointer");
    if (matrix->rows == 0) return 0;
    if (!p || !dy || (matrix->cols > 0 && !dx)) return skg_fail("null device CSR or vector pointer");
    if ((rc = skk_spmv_f64(&p->A.k, dx, dy))) return skg_hip_fail("CSR SpMV launch", rc);
    return 0;
}

int sk_gpu_spmv_transpose_device_f64(const sk_gpu_csr *matrix, const double *dy, double *dx)
{                                                  /* cuda/src/sankhya_cuda.cu:172-190 */
    const csr_plan *p = plan_of(matrix);
    int rc;
    skg_clear_error();
    if (!matrix) return skg_fail("null device CSR pointer");
    if (matrix->rows == 0 || matrix->cols == 0) return 0;
    if (!p || !dy || !dx) return skg_fail("null device CSR or transpose vector pointer");
    if ((rc = skk_spmv_f64(&p->AT.k, dy, dx))) return skg_hip_fail("transpose SpMV launch", rc);
    return 0;
}

int sk_gpu_axpy_device_f64(int n, double alpha, const double *dx, double *dy)
{                                                  /* cuda/src/sankhya_cuda.cu:192-197 */
    int rc;
    skg_clear_error();
    if (n < 0) return skg_fail("negative vector dimension");
    if (n == 0) return 0;
    if (!dx || !dy) return skg_fail("null device vector pointer");
    if ((rc = skk_axpy_f64(n, alpha, dx, dy))) return skg_hip_fail("AXPY launch", rc);
    return 0;
}

int sk_gpu_synchronize(void)
{
    const int rc = skk_sync();
    return rc ? skg_hip_fail("stream synchronize", rc) : 0;
}

int sk_gpu_spmv_f64(int rows, int cols, int nnz, const int *row_offsets, const int *column_indices,
                    const double *values, const double *x, double *y)
{                                                  /* cuda/src/sankhya_cuda.cu:199-252 */
    sk_gpu_csr mat;
    double *dx = NULL, *dy = NULL;
    int rc;
    skg_clear_error();
    if (rows < 0 || cols < 0 || nnz < 0) return skg_fail("negative CSR dimension");
    if (!row_offsets || !y || (nnz > 0 && (!column_indices || !values)) || (cols > 0 && !x))
        return skg_fail("null CSR or vector pointer");
    if (rows == 0) return 0;
    memset(&mat, 0, sizeof(mat));
    if (sk_gpu_csr_create(&mat, rows, cols, nnz, row_offsets, column_indices, values)) return -1;
    if ((rc = skk_malloc((void **)&dx, sizeof(double) * (size_t)(cols ? cols : 1))) ||
        (rc = skk_malloc((void **)&dy, sizeof(double) * (size_t)rows)) ||
        (rc = skk_h2d(dx, x, sizeof(double) * (size_t)cols))) {
        skk_free(dx); skk_free(dy); sk_gpu_csr_destroy(&mat);
        return skg_hip_fail("SpMV staging", rc);
    }
    rc = sk_gpu_spmv_device_f64(&mat, dx, dy);
    if (!rc && (rc = skk_d2h(y, dy, sizeof(double) * (size_t)rows))) rc = skg_hip_fail("cudaMemcpy y", rc);
    skk_free(dx); skk_free(dy); sk_gpu_csr_destroy(&mat);
    return rc;
}

/* The mixed engine's FP32 product on host arrays: the FP64 values are
 * rounded to FP32 once, x is FP32, products and sums are FP32 (the kernel
 * of the FP32 iterations).  transpose = 1 multiplies by A^T through the
 * stored CSR(A^T), as the engine does. */
int sk_gpu_spmv_f32(int rows, int cols, int nnz, const int *row_offsets, const int *column_indices,
                    const double *values, int transpose, const float *x, float *y)
{
    hcsr a, at;
    dcsr d;
    float *dx = NULL, *dy = NULL;
    int rc, in, out;
    skg_clear_error();
    if (rows < 0 || cols < 0 || nnz < 0) return skg_fail("negative CSR dimension");
    if (!row_offsets || !y || (nnz > 0 && (!column_indices || !values)) || !x)
        return skg_fail("null CSR or vector pointer");
    if (hcsr_validate(rows, cols, nnz, row_offsets, column_indices, values)) return -1;
    in = transpose ? rows : cols;
    out = transpose ? cols : rows;
    if (out == 0) return 0;
    if (skg_have_device()) return -1;
    a.rows = rows; a.cols = cols; a.nnz = nnz;
    a.p = (int *)row_offsets; a.j = (int *)column_indices; a.v = (double *)values;
    memset(&at, 0, sizeof(at));
    if (transpose && hcsr_transpose(&a, &at)) return skg_fail("out of host memory");
    rc = dcsr_upload(&d, transpose ? &at : &a, 1);
    hcsr_free(&at);
    if (rc) return -1;
    if ((rc = skk_malloc((void **)&dx, sizeof(float) * (size_t)(in ? in : 1))) ||
        (rc = skk_malloc((void **)&dy, sizeof(float) * (size_t)out)) ||
        (rc = skk_h2d(dx, x, sizeof(float) * (size_t)in))) {
        skk_free(dx); skk_free(dy); dcsr_free(&d);
        return skg_hip_fail("SpMV staging", rc);
    }
    if (d.k.ntiles == 0) {   /* no rows to tile: the product is zero */
        memset(y, 0, sizeof(float) * (size_t)out);
        skk_free(dx); skk_free(dy); dcsr_free(&d);
        return 0;
    }
    if ((rc = skk_spmv_f32(&d.k, dx, dy))) rc = skg_hip_fail("FP32 SpMV launch", rc);
    if (!rc && (rc = skk_d2h(y, dy, sizeof(float) * (size_t)out))) rc = skg_hip_fail("copy y", rc);
    skk_free(dx); skk_free(dy); dcsr_free(&d);
    return rc;
}

int sk_gpu_axpy_f64(int n, double alpha, const double *x, double *y)
{                                                  /* cuda/src/sankhya_cuda.cu:254-284 */
    double *dx = NULL, *dy = NULL;
    int rc;
    skg_clear_error();
    if (n < 0) return skg_fail("negative vector dimension");
    if (n == 0) return 0;
    if (!x || !y) return skg_fail("null vector pointer");
    if (skg_have_device()) return -1;
    if ((rc = skk_malloc((void **)&dx, sizeof(double) * (size_t)n)) ||
        (rc = skk_malloc((void **)&dy, sizeof(double) * (size_t)n)) ||
        (rc = skk_h2d(dx, x, sizeof(double) * (size_t)n)) ||
        (rc = skk_h2d(dy, y, sizeof(double) * (size_t)n)) ||
        (rc = skk_axpy_f64(n, alpha, dx, dy)) ||
        (rc = skk_d2h(y, dy, sizeof(double) * (size_t)n))) {
        skk_free(dx); skk_free(dy);
        return skg_hip_fail("AXPY", rc);
    }
    skk_free(dx); skk_free(dy);
    return 0;
}

/* ----------------------------------------------- plain PDHG (legacy) */

static int valid_lower(double v) { return !isnan(v) && !(isinf(v) && v > 0.0); }
static int valid_upper(double v) { return !isnan(v) && !(isinf(v) && v < 0.0); }

static void result_reset(sk_gpu_lp_result *r)
{
    r->status = -1;
    r->iterations = 0;
    r->objective = NAN;
    r->maximum_row_violation = r->maximum_step = r->maximum_kkt_residual = INFINITY;
    r->maximum_dual_residual = r->maximum_complementarity = INFINITY;
}

/* cuda/src/sankhya_cuda_lp.cu:101-385 (solve_qp): same input contract,
 * status codes and checkpoint tests; the iteration runs as two fused
 * product kernels (dual prox in A xbar, primal prox in A^T y) and the
 * checkpoint residuals are reduced on the device. */
static int solve_qp(const sk_gpu_csr *matrix, const sk_gpu_csr *hessian, const double *qd,
                    const double *c, const double *rl, const double *ru, const double *cl,
                    const double *cu, sk_gpu_lp_settings st, const double *ps, const double *ds,
                    double *solution, double *dual_solution, sk_gpu_lp_result *res)
{
    const csr_plan *pa, *ph;
    skk_legacy L;
    void *bufs[20];
    int nb = 0, rows, cols, i, j, rc = 0, it, final_iteration, have_prev = 0, cap, k;
    double *h_out = NULL, start;
    skg_clear_error();
    if (!res) return skg_fail("null result");
    result_reset(res);
    if (skg_have_device()) return -1;
    if (!matrix || !c || !rl || !ru || !cl || !cu || !solution) return skg_fail("null input");
    if (st.max_iterations <= 0 || st.check_every <= 0 || st.theta < 0.0 || st.theta > 1.0 ||
        st.tolerance <= 0.0 || st.time_limit < 0.0) return skg_fail("invalid settings");
    if ((ps == NULL) != (ds == NULL)) return skg_fail("step vectors must be given together");
    if (ps == NULL && (st.tau <= 0.0 || st.sigma <= 0.0)) return skg_fail("non-positive scalar step");
    rows = matrix->rows; cols = matrix->cols;
    if (rows < 0 || cols < 0) return skg_fail("negative dimension");
    for (j = 0; j < cols; ++j)
        if (!isfinite(c[j]) || !valid_lower(cl[j]) || !valid_upper(cu[j]) || cl[j] > cu[j])
            return skg_fail("invalid cost or column bound");
    for (i = 0; i < rows; ++i)
        if (!valid_lower(rl[i]) || !valid_upper(ru[i]) || rl[i] > ru[i]) return skg_fail("invalid row bound");
    if (ps) {
        for (j = 0; j < cols; ++j) if (!isfinite(ps[j]) || ps[j] <= 0.0) return skg_fail("invalid primal step");
        for (i = 0; i < rows; ++i) if (!isfinite(ds[i]) || ds[i] <= 0.0) return skg_fail("invalid dual step");
    }
    if (hessian && (qd || hessian->rows != cols || hessian->cols != cols)) return skg_fail("invalid Hessian");
    if (qd) for (j = 0; j < cols; ++j) if (!isfinite(qd[j]) || qd[j] < 0.0) return skg_fail("invalid diagonal");
    pa = plan_of(matrix);
    ph = plan_of(hessian);
    if (!pa || (hessian && !ph)) return skg_fail("CSR handle was not created");

    memset(&L, 0, sizeof(L));
    L.m = rows; L.n = cols;
    L.A = &pa->A.k; L.AT = &pa->AT.k; L.Q = hessian ? &ph->A.k : NULL;
    L.tau = st.tau; L.sigma = st.sigma; L.theta = st.theta;
    cap = pa->A.k.grid;
    if (pa->AT.k.grid > cap) cap = pa->AT.k.grid;
    k = ((rows > cols ? rows : cols) + SKK_BLOCK - 1) / SKK_BLOCK;
    if (k < 1) k = 1;
    if (k > 1024) k = 1024;
    L.grid = k;
#define DALLOC(ptr, count)                                                              \
    do {                                                                                \
        if ((rc = skk_malloc((void **)&(ptr), sizeof(double) * (size_t)((count) > 0 ? (count) : 1)))) \
            goto fail;                                                                  \
        bufs[nb++] = (void *)(ptr);                                                     \
    } while (0)
#define DUP(dst, src, count)                                                            \
    do {                                                                                \
        double *tmp_;                                                                   \
        DALLOC(tmp_, count);                                                            \
        if ((rc = skk_h2d(tmp_, (src), sizeof(double) * (size_t)(count)))) goto fail;   \
        (dst) = tmp_;                                                                   \
    } while (0)
    DALLOC(L.x, cols); DALLOC(L.xbar, cols); DALLOC(L.grad, cols); DALLOC(L.prev, cols);
    DALLOC(L.y, rows); DALLOC(L.act, rows);
    if (hessian) DALLOC(L.qg, cols);
    DUP(L.c, c, cols); DUP(L.l, cl, cols); DUP(L.u, cu, cols); DUP(L.rl, rl, rows); DUP(L.ru, ru, rows);
    if (ps) { DUP(L.tau_vec, ps, cols); DUP(L.sigma_vec, ds, rows); }
    if (qd) DUP(L.qd, qd, cols);
    DALLOC(L.part, 8 * L.grid);
    DALLOC(L.out, 8);
    if ((rc = skk_host_alloc((void **)&h_out, sizeof(double) * 8))) goto fail;
    if ((rc = skk_legacy_init(L.x, L.xbar, L.l, L.u, cols, L.y, rows))) goto fail;
    final_iteration = st.max_iterations;
    start = skg_now();
    for (it = 1; it <= st.max_iterations; ++it) {
        if (st.time_limit > 0.0 && it % st.check_every == 1 && skg_now() - start >= st.time_limit) {
            final_iteration = it - 1; res->status = 2; break;
        }
        if ((rc = skk_legacy_iter(&L))) goto fail;
        if (it % st.check_every == 0 || it == st.max_iterations) {
            double objective, step, scale, viol, kkt, dres, comp;
            if ((rc = skk_legacy_check(&L, have_prev)) ||
                (rc = skk_d2h_async(h_out, L.out, sizeof(double) * 8)) || (rc = skk_sync())) goto fail;
            objective = h_out[0];
            if (h_out[1] > 0.0 || !isfinite(objective)) {
                res->status = -2;
                for (k = 0; k < nb; ++k) skk_free(bufs[k]);
                skk_host_free(h_out);
                skg_fail("non-finite PDHG iterate");
                return -2;
            }
            step = have_prev ? h_out[2] : INFINITY;
            have_prev = 1;
            scale = h_out[3];
            viol = h_out[4] > 0.0 ? h_out[4] : 0.0;
            kkt = h_out[5]; dres = h_out[6]; comp = h_out[7];
            res->objective = objective;
            res->maximum_row_violation = viol;
            res->maximum_step = step;
            res->maximum_kkt_residual = kkt;
            res->maximum_dual_residual = dres;
            res->maximum_complementarity = comp / (1.0 + fabs(objective));
            if (viol <= st.tolerance && step <= st.tolerance * scale && kkt <= st.tolerance &&
                dres <= st.tolerance && res->maximum_complementarity <= st.tolerance) {
                final_iteration = it; res->status = 0; break;
            }
        }
    }
    if (res->status < 0) res->status = 1;
    res->iterations = final_iteration;
    if ((rc = skk_d2h(solution, L.x, sizeof(double) * (size_t)cols))) goto fail;
    if (dual_solution && rows > 0 && (rc = skk_d2h(dual_solution, L.y, sizeof(double) * (size_t)rows))) goto fail;
    for (k = 0; k < nb; ++k) skk_free(bufs[k]);
    skk_host_free(h_out);
    return 0;
fail:
    for (k = 0; k < nb; ++k) skk_free(bufs[k]);
    skk_host_free(h_out);
    return skg_hip_fail("PDHG device operation", rc) > 0 ? -1 : -1;
#undef DALLOC
#undef DUP
}

int sk_gpu_lp_pdhg(const sk_gpu_csr *matrix, const double *c, const double *row_lower,
                   const double *row_upper, const double *col_lower, const double *col_upper,
                   sk_gpu_lp_settings settings, double *solution, sk_gpu_lp_result *result)
{
    return solve_qp(matrix, NULL, NULL, c, row_lower, row_upper, col_lower, col_upper, settings,
                    NULL, NULL, solution, NULL, result);
}

int sk_gpu_diagonal_qp_pdhg(const sk_gpu_csr *matrix, const double *qd, const double *c,
                            const double *row_lower, const double *row_upper, const double *col_lower,
                            const double *col_upper, sk_gpu_lp_settings settings, double *solution,
                            sk_gpu_lp_result *result)
{
    if (!qd) return skg_fail("null diagonal");
    return solve_qp(matrix, NULL, qd, c, row_lower, row_upper, col_lower, col_upper, settings, NULL,
                    NULL, solution, NULL, result);
}

int sk_gpu_sparse_qp_pdhg(const sk_gpu_csr *matrix, const sk_gpu_csr *hessian, const double *c,
                          const double *row_lower, const double *row_upper, const double *col_lower,
                          const double *col_upper, sk_gpu_lp_settings settings, double *solution,
                          sk_gpu_lp_result *result)
{
    if (!hessian) return skg_fail("null Hessian");
    return solve_qp(matrix, hessian, NULL, c, row_lower, row_upper, col_lower, col_upper, settings,
                    NULL, NULL, solution, NULL, result);
}

int sk_gpu_qp_pdhg_preconditioned(const sk_gpu_csr *matrix, const sk_gpu_csr *hessian,
                                  const double *qd, const double *c, const double *row_lower,
                                  const double *row_upper, const double *col_lower,
                                  const double *col_upper, const double *primal_steps,
                                  const double *dual_steps, sk_gpu_lp_settings settings,
                                  double *solution, sk_gpu_lp_result *result)
{
    if (!primal_steps || !dual_steps) return skg_fail("null step vectors");
    return solve_qp(matrix, hessian, qd, c, row_lower, row_upper, col_lower, col_upper, settings,
                    primal_steps, dual_steps, solution, NULL, result);
}

int sk_gpu_qp_pdhg_preconditioned_with_dual(const sk_gpu_csr *matrix, const sk_gpu_csr *hessian,
                                            const double *qd, const double *c,
                                            const double *row_lower, const double *row_upper,
                                            const double *col_lower, const double *col_upper,
                                            const double *primal_steps, const double *dual_steps,
                                            sk_gpu_lp_settings settings, double *solution,
                                            double *dual_solution, sk_gpu_lp_result *result)
{
    if (!primal_steps || !dual_steps || !dual_solution) return skg_fail("null step or dual vectors");
    return solve_qp(matrix, hessian, qd, c, row_lower, row_upper, col_lower, col_upper, settings,
                    primal_steps, dual_steps, solution, dual_solution, result);
}

/* ------------------------------------------------------ certificates */

typedef struct vws {         /* device workspace of one certificate call */
    dcsr A, AT, Q;
    double *c, *rl, *ru, *l, *u, *x, *y, *act, *rc, *qx, *part, *out;
    skk_verify_in V;
} vws;

static void vws_free(vws *w)
{
    dcsr_free(&w->A); dcsr_free(&w->AT); dcsr_free(&w->Q);
    skk_free(w->c); skk_free(w->rl); skk_free(w->ru); skk_free(w->l); skk_free(w->u);
    skk_free(w->x); skk_free(w->y); skk_free(w->act); skk_free(w->rc); skk_free(w->qx);
    skk_free(w->part); skk_free(w->out);
    memset(w, 0, sizeof(*w));
}

static int dup_d(double **dst, const double *src, int n)
{
    int rc = skk_malloc((void **)dst, sizeof(double) * (size_t)(n ? n : 1));
    if (!rc && src && n) rc = skk_h2d(*dst, src, sizeof(double) * (size_t)n);
    return rc;
}

/* Upload CSR(A), CSR(A^T) (and Q) plus the vectors for a certificate. */
static int vws_build(vws *w, const hcsr *A, const hcsr *Qh, const double *c, const double *rl,
                     const double *ru, const double *l, const double *u, const double *x,
                     const double *y)
{
    hcsr at;
    int rc, g, cap;
    const int m = A->rows, n = A->cols;
    memset(w, 0, sizeof(*w));
    if (hcsr_transpose(A, &at)) return skg_fail("out of host memory");
    rc = plan_upload(&w->A, A);
    if (!rc) rc = plan_upload(&w->AT, &at);
    if (!rc && Qh) rc = plan_upload(&w->Q, Qh);
    hcsr_free(&at);
    if (rc) { vws_free(w); return -1; }
    g = ((m > n ? m : n) + SKK_BLOCK - 1) / SKK_BLOCK;
    if (g < 1) g = 1;
    if (g > 1024) g = 1024;
    cap = g;
    if ((rc = dup_d(&w->c, c, n)) || (rc = dup_d(&w->rl, rl, m)) || (rc = dup_d(&w->ru, ru, m)) ||
        (rc = dup_d(&w->l, l, n)) || (rc = dup_d(&w->u, u, n)) || (rc = dup_d(&w->x, x, n)) ||
        (y && (rc = dup_d(&w->y, y, m))) || (rc = dup_d(&w->act, NULL, m)) ||
        (rc = dup_d(&w->rc, NULL, n)) || (rc = dup_d(&w->qx, NULL, n)) ||
        (rc = dup_d(&w->part, NULL, 8 * cap)) || (rc = dup_d(&w->out, NULL, 8))) {
        vws_free(w);
        return skg_hip_fail("certificate upload", rc);
    }
    w->V.m = m; w->V.n = n;
    w->V.A = w->A.k; w->V.AT = w->AT.k; w->V.Q = Qh ? &w->Q.k : NULL;
    w->V.c = w->c; w->V.rl = w->rl; w->V.ru = w->ru; w->V.l = w->l; w->V.u = w->u;
    w->V.x = w->x; w->V.y = y ? w->y : NULL;
    w->V.act = w->act; w->V.rc = w->rc; w->V.qx = w->qx;
    w->V.part = w->part; w->V.out = w->out; w->V.grid = g;
    return 0;
}

/* sk_verify (c/src/sk_core.c:394-510) evaluated on the device in FP64. */
sk_status sk_gpu_verify(const sk_model *m, sk_solution *s)
{
    hcsr A, Q;
    vws w;
    double out[8];
    int rc;
    skg_clear_error();
    if (!m || !s || !s->x) return SK_ERR_ARG;
    if (sk_gpu_model_validate(m) != SK_OK) return SK_ERR_STRUCTURE;
    if (skg_have_device()) return SK_ERR_UNSUPPORTED;
    if (hcsr_from_csc(&m->A, &A)) return SK_ERR_MEMORY;
    memset(&Q, 0, sizeof(Q));
    if (m->Q) {   /* full symmetric: CSC arrays read as CSR are Q itself */
        Q.rows = Q.cols = m->ncol; Q.nnz = m->Q->p[m->ncol];
        Q.p = m->Q->p; Q.j = m->Q->i; Q.v = m->Q->x;
    }
    if (vws_build(&w, &A, m->Q ? &Q : NULL, m->c, m->rlow, m->rupp, m->clow, m->cupp, s->x, s->y)) {
        hcsr_free(&A);
        return SK_ERR_MEMORY;
    }
    hcsr_free(&A);
    if ((rc = skk_verify(&w.V)) || (rc = skk_d2h(out, w.out, sizeof(double) * 6))) {
        skg_hip_fail("certificate", rc); vws_free(&w); return SK_ERR_NUMERIC;
    }
    if (s->rowact && m->nrow) skk_d2h(s->rowact, w.act, sizeof(double) * (size_t)m->nrow);
    if (s->rc && s->y && m->ncol) skk_d2h(s->rc, w.rc, sizeof(double) * (size_t)m->ncol);
    vws_free(&w);
    {
        /* out: c'x, dual-objective terms, x'Qx, pinf, dinf, max complementarity product */
        const double pobj = m->objshift + out[0] + (m->Q ? 0.5 * out[2] : 0.0);
        const double dobj = m->objshift + out[1];
        s->objective = pobj;
        s->primal_infeasibility = out[3] > 0.0 ? out[3] : 0.0;
        if (s->y) {
            s->dual_infeasibility = out[4] > 0.0 ? out[4] : 0.0;
            s->complementarity = m->Q ? out[5] / (1.0 + fabs(pobj)) : fabs(pobj - dobj) / (1.0 + fabs(pobj));
        } else {
            s->dual_infeasibility = INFINITY;
            s->complementarity = INFINITY;
        }
    }
    return SK_OK;
}

static double finite_norm_sentinel(const double *v, int n)
{
    double acc = 0.0;
    int k;
    for (k = 0; k < n; ++k) if (isfinite(v[k]) && fabs(v[k]) < SK_INFINITY) acc += v[k] * v[k];
    return sqrt(acc);
}

static int relkkt_hcsr(const hcsr *A, const hcsr *Q, const double *c, const double *rl,
                       const double *ru, const double *l, const double *u, const double *x,
                       const double *y, double out[5])
{
    vws w;
    double s[5];
    int rc;
    const int m = A->rows, n = A->cols;
    double cost_norm, bound_norm;
    if (vws_build(&w, A, Q, c, rl, ru, l, u, x, y)) return -1;
    /* act = A x and rc = c + Qx + A^T y come from the certificate pass */
    if ((rc = skk_verify(&w.V)) || (rc = skk_relative_kkt(&w.V)) ||
        (rc = skk_d2h(s, w.out, sizeof(double) * 5))) {
        vws_free(&w);
        return skg_hip_fail("relative KKT", rc);
    }
    vws_free(&w);
    cost_norm = finite_norm_sentinel(c, n);
    bound_norm = sqrt(finite_norm_sentinel(rl, m) * finite_norm_sentinel(rl, m) +
                      finite_norm_sentinel(ru, m) * finite_norm_sentinel(ru, m));
    out[0] = sqrt(fmax(0.0, s[0])) / (1.0 + bound_norm);
    out[1] = sqrt(fmax(0.0, s[1])) / (1.0 + cost_norm);
    out[3] = s[2] + 0.5 * s[4];            /* c'x + 1/2 x'Qx              */
    out[4] = s[3] - 0.5 * s[4];            /* sum of terms - 1/2 x'Qx     */
    out[2] = fabs(out[3] - out[4]) / (1.0 + fabs(out[3]) + fabs(out[4]));
    if (!isfinite(out[0]) || !isfinite(out[1]) || !isfinite(out[2])) out[0] = out[1] = out[2] = INFINITY;
    return 0;
}

int sk_gpu_relative_kkt(int rows, int cols, int nnz, const int *row_offsets, const int *column_indices,
                        const double *values, const double *c, const double *row_lower,
                        const double *row_upper, const double *col_lower, const double *col_upper,
                        const double *x, const double *y, double out[5])
{
    hcsr A;
    skg_clear_error();
    if (!c || !row_lower || !row_upper || !col_lower || !col_upper || !x || (rows && !y) || !out)
        return skg_fail("null input");
    if (hcsr_validate(rows, cols, nnz, row_offsets, column_indices, values)) return -1;
    if (skg_have_device()) return -1;
    A.rows = rows; A.cols = cols; A.nnz = nnz;
    A.p = (int *)row_offsets; A.j = (int *)column_indices; A.v = (double *)values;
    return relkkt_hcsr(&A, NULL, c, row_lower, row_upper, col_lower, col_upper, x, y, out);
}

int sk_gpu_relative_kkt_qp(int rows, int cols, int nnz, const int *row_offsets, const int *column_indices,
                           const double *values, const int *q_row_offsets, const int *q_column_indices,
                           const double *q_values, int q_nnz, const double *c, const double *row_lower,
                           const double *row_upper, const double *col_lower, const double *col_upper,
                           const double *x, const double *y, double out[5])
{
    hcsr A, Q;
    skg_clear_error();
    if (!c || !row_lower || !row_upper || !col_lower || !col_upper || !x || (rows && !y) || !out)
        return skg_fail("null input");
    if (hcsr_validate(rows, cols, nnz, row_offsets, column_indices, values)) return -1;
    if (q_row_offsets && hcsr_validate(cols, cols, q_nnz, q_row_offsets, q_column_indices, q_values)) return -1;
    if (skg_have_device()) return -1;
    A.rows = rows; A.cols = cols; A.nnz = nnz;
    A.p = (int *)row_offsets; A.j = (int *)column_indices; A.v = (double *)values;
    Q.rows = Q.cols = cols; Q.nnz = q_nnz;
    Q.p = (int *)q_row_offsets; Q.j = (int *)q_column_indices; Q.v = (double *)q_values;
    return relkkt_hcsr(&A, q_row_offsets ? &Q : NULL, c, row_lower, row_upper, col_lower, col_upper, x, y, out);
}

int sk_gpu_relative_kkt_device_noalloc(const sk_gpu_csr *matrix, const double *device_c,
                                       const double *device_row_lower, const double *device_row_upper,
                                       const double *device_col_lower, const double *device_col_upper,
                                       const double *device_x, const double *device_y, double *device_out)
{
    const csr_plan *p = plan_of(matrix);
    int rc;
    skg_clear_error();
    if (!matrix || !p) return skg_fail("null device CSR");
    if (!device_out || (matrix->cols > 0 && (!device_c || !device_col_lower || !device_col_upper || !device_x)) ||
        (matrix->rows > 0 && (!device_row_lower || !device_row_upper || !device_y)))
        return skg_fail("null device vector pointer");
    if (p->A.k.grid > SKK_SPMV_GRID || p->AT.k.grid > SKK_SPMV_GRID)
        return skg_fail("product grid of %d / %d workgroups exceeds the %d partials of the static scratch",
                        p->A.k.grid, p->AT.k.grid, SKK_SPMV_GRID);
    if ((rc = skk_relkkt_noalloc(&p->A.k, &p->AT.k, device_c, device_row_lower, device_row_upper, device_col_lower,
                                 device_col_upper, device_x, device_y, device_out)))
        return skg_hip_fail("relative KKT (no allocation)", rc);
    return 0;
}

/* -------------------------------------------------- sk_solve mirror */

void sk_gpu_solver_options_default(sk_gpu_options *g)
{
    memset(g, 0, sizeof(*g));
    sk_gpu_pdlp_params_default(&g->pdlp);
}

/* c/src/sk_solve.c:97-132 */
static int interval_infeasible(const sk_model *m, double tol)
{
    int i, j, p, impossible = 0;
    double *mn, *mx;
    unsigned char *mni, *mxi;
    for (j = 0; j < m->ncol; ++j) if (m->clow[j] > m->cupp[j] + tol) return 1;
    mn = (double *)calloc((size_t)m->nrow + 1, sizeof(double));
    mx = (double *)calloc((size_t)m->nrow + 1, sizeof(double));
    mni = (unsigned char *)calloc((size_t)m->nrow + 1, 1);
    mxi = (unsigned char *)calloc((size_t)m->nrow + 1, 1);
    if (!mn || !mx || !mni || !mxi) { free(mn); free(mx); free(mni); free(mxi); return 0; }
    for (j = 0; j < m->ncol; ++j)
        for (p = m->A.p[j]; p < m->A.p[j + 1]; ++p) {
            const int row = m->A.i[p];
            const double v = m->A.x[p];
            const double lo = v >= 0.0 ? m->clow[j] : m->cupp[j];
            const double hi = v >= 0.0 ? m->cupp[j] : m->clow[j];
            if ((v >= 0.0 && SK_IS_NEG_INF(lo)) || (v < 0.0 && SK_IS_INF(lo))) mni[row] = 1;
            else if (!mni[row]) mn[row] += v * lo;
            if ((v >= 0.0 && SK_IS_INF(hi)) || (v < 0.0 && SK_IS_NEG_INF(hi))) mxi[row] = 1;
            else if (!mxi[row]) mx[row] += v * hi;
        }
    for (i = 0; i < m->nrow; ++i) {
        if (!mxi[i] && m->rlow[i] > mx[i] + tol) { impossible = 1; break; }
        if (!mni[i] && m->rupp[i] < mn[i] - tol) { impossible = 1; break; }
    }
    free(mn); free(mx); free(mni); free(mxi);
    return impossible;
}

/* The certificate margins of the reference's first-order path
 * (c/src/sk_solve.c:987-996): a QP needs sk_verify's primal infeasibility
 * within 100 primal_tol and its dual infeasibility and complementarity
 * within 100 dual_tol.  An LP there stops only with its largest row
 * violation within primal_tol (sk_solve.c:947); this engine stops on the
 * relative KKT instead, so an LP needs the verified primal infeasibility
 * within the same 100 primal_tol margin (its dual sign test, with the 1e-9
 * "at bound" rule, is not an LP first-order certificate: sk_solve.c:983-985). */
static int certificate_ok(const sk_model *m, const sk_options *o, const sk_solution *s)
{
    const double pt = o->primal_tol > 0.0 ? o->primal_tol : 1e-7;
    const double dt = o->dual_tol > 0.0 ? o->dual_tol : 1e-7;
    if (!(s->primal_infeasibility <= 100.0 * pt)) return 0;
    if (m->Q && !(s->dual_infeasibility <= 100.0 * dt && s->complementarity <= 100.0 * dt)) return 0;
    return 1;
}

/* sk_solve (c/src/sk_solve.c:1009-1046) for continuous LP / convex QP on the
 * GPU PDLP engine.  OPTIMAL needs both the relative KKT triple of the
 * ORIGINAL model and the sk_verify certificate (certificate_ok), each
 * evaluated on the device in FP64.  When the scaled loop converges but
 * either is not yet met, the solve continues from the same point with a 10x
 * tighter loop tolerance (at most three times), within what is left of the
 * iteration and time limits.  Otherwise the result is ITERATION_LIMIT or
 * TIME_LIMIT, as the reference reports an unconverged first-order solve
 * (sk_solve.c:997-999); NUMERIC_FAILURE for a non-finite result. */
sk_status sk_gpu_solve(const sk_model *m, const sk_options *o, const sk_gpu_options *g, sk_solution *s,
                       sk_gpu_pdlp_stats *stats)
{
    sk_options od;
    sk_gpu_options gd;
    sk_gpu_pdlp_stats acc;
    hcsr A, Qh;
    double *x, *y, kkt[5], t0 = skg_now(), eps;
    int j, rc, round, verified = 0, cert = 0;
    long long iters = 0;
    skg_clear_error();
    memset(&acc, 0, sizeof(acc));
    if (stats) memset(stats, 0, sizeof(*stats));
    if (!m || !s) return SK_ERR_ARG;
    if (sk_gpu_model_validate(m) != SK_OK) return SK_ERR_STRUCTURE;
    if (!o) { sk_gpu_options_default(&od); o = &od; }
    if (!g) { sk_gpu_solver_options_default(&gd); g = &gd; }
    if (o->lp_engine < SK_LP_AUTO || o->lp_engine > SK_LP_FIRST_ORDER) return SK_ERR_ARG;
    if (interval_infeasible(m, o->primal_tol > 0.0 ? o->primal_tol : 1e-7)) {
        sk_gpu_solution_init(s);
        s->result = SK_RESULT_INFEASIBLE; s->ncol = m->ncol; s->nrow = m->nrow;
        s->primal_infeasibility = INFINITY;
        return SK_OK;
    }
    for (j = 0; j < m->ncol; ++j) if (m->vartype[j] == SK_INTEGER) return SK_ERR_UNSUPPORTED;
    if (o->lp_engine == SK_LP_SIMPLEX) return SK_ERR_UNSUPPORTED;
    if (hcsr_from_csc(&m->A, &A)) return SK_ERR_MEMORY;
    memset(&Qh, 0, sizeof(Qh));
    if (m->Q) { Qh.rows = Qh.cols = m->ncol; Qh.nnz = m->Q->p[m->ncol]; Qh.p = m->Q->p; Qh.j = m->Q->i; Qh.v = m->Q->x; }
    x = (double *)calloc((size_t)m->ncol + 1, sizeof(double));
    y = (double *)calloc((size_t)m->nrow + 1, sizeof(double));
    sk_gpu_solution_init(s);
    s->rowact = (double *)calloc((size_t)m->nrow + 1, sizeof(double));
    s->rc = (double *)calloc((size_t)m->ncol + 1, sizeof(double));
    if (!x || !y || !s->rowact || !s->rc) {
        free(x); free(y); free(s->rowact); free(s->rc); s->rowact = s->rc = NULL;
        hcsr_free(&A); return SK_ERR_MEMORY;
    }
    s->x = x; s->y = y;
    s->ncol = m->ncol; s->nrow = m->nrow;
    {
        sk_gpu_pdlp_params P = g->pdlp;
        sk_gpu_pdlp_session *S = NULL;
        long long budget;
        if (o->iteration_limit > 0 && o->iteration_limit < P.max_iterations) P.max_iterations = (int)o->iteration_limit;
        budget = P.max_iterations;
        if (o->time_limit > 0.0) P.time_limit_seconds = o->time_limit - (skg_now() - t0);
        eps = P.eps_optimal;
        kkt[0] = kkt[1] = kkt[2] = INFINITY;
        /* the full symmetric Q's CSC arrays read as CSR are Q itself */
        if (sk_gpu_pdlp_session_create_qp(m->nrow, m->ncol, A.nnz, A.p, A.j, A.v, m->Q ? m->Q->p : NULL,
                                          m->Q ? m->Q->i : NULL, m->Q ? m->Q->x : NULL,
                                          m->Q ? m->Q->p[m->ncol] : 0, m->c, m->rlow, m->rupp, m->clow,
                                          m->cupp, &P, &S)) {
            hcsr_free(&A); sk_gpu_solution_free(s); return SK_ERR_NUMERIC;
        }
        for (round = 0; round < 4; ++round) {
            if (round > 0) {
                /* continue the same loop (iterate, step, primal weight,
                   averages) with a 10x tighter tolerance, within what is
                   left of the iteration and time limits */
                const double left = o->time_limit > 0.0 ? o->time_limit - (skg_now() - t0) : 0.0;
                const long long left_it = budget - acc.iterations;
                if (left_it <= 0 || (o->time_limit > 0.0 && left <= 0.0)) break;
                P.eps_optimal *= 0.1;
                if (sk_gpu_pdlp_session_set_limits(S, (int)left_it, left, P.eps_optimal)) break;
                rc = sk_gpu_pdlp_session_resume(S, x, y, &acc);
            } else {
                if (o->time_limit > 0.0 && P.time_limit_seconds <= 0.0) break;
                rc = sk_gpu_pdlp_session_solve(S, g->x0, g->y0, x, y, &acc);
            }
            if (rc == -1) { sk_gpu_pdlp_session_destroy(S); hcsr_free(&A); sk_gpu_solution_free(s); return SK_ERR_NUMERIC; }
            iters = acc.iterations;
            verified = 0;
            if (rc == -2) break;
            if (relkkt_hcsr(&A, m->Q ? &Qh : NULL, m->c, m->rlow, m->rupp, m->clow, m->cupp, x, y, kkt)) {
                sk_gpu_pdlp_session_destroy(S); hcsr_free(&A); sk_gpu_solution_free(s); return SK_ERR_NUMERIC;
            }
            if ((P.verbosity > 0 || o->verbosity > 0) && (!acc.converged || acc.cancelled))
                fprintf(stderr, "sk_gpu_solve round %d: eps %.1e, %d iterations, not converged (rc %d, cancelled "
                        "%d), relative KKT %.3e %.3e %.3e\n", round, P.eps_optimal, acc.iterations, rc,
                        acc.cancelled, kkt[0], kkt[1], kkt[2]);
            if (!acc.converged || acc.cancelled) break;
            if (sk_gpu_verify(m, s) != SK_OK) break;
            verified = 1;
            cert = certificate_ok(m, o, s);
            if (P.verbosity > 0 || o->verbosity > 0)
                fprintf(stderr, "sk_gpu_solve round %d: eps %.1e, %d iterations, relative KKT %.3e %.3e %.3e, "
                        "verify pinf %.3e dinf %.3e comp %.3e -> %s\n", round, P.eps_optimal, acc.iterations,
                        kkt[0], kkt[1], kkt[2], s->primal_infeasibility, s->dual_infeasibility,
                        s->complementarity, cert ? "certified" : "not certified");
            if (kkt[0] <= eps && kkt[1] <= eps && kkt[2] <= eps && cert) break;
        }
        sk_gpu_pdlp_session_destroy(S);
    }
    hcsr_free(&A);
    s->iterations = iters;
    if (!verified && sk_gpu_verify(m, s) != SK_OK) {
        s->result = SK_RESULT_NUMERIC_FAILURE;
        if (stats) *stats = acc;
        return SK_OK;
    }
    cert = certificate_ok(m, o, s);
    s->solve_seconds = skg_now() - t0;
    if (!isfinite(s->objective) || !isfinite(s->primal_infeasibility)) s->result = SK_RESULT_NUMERIC_FAILURE;
    else if (kkt[0] <= eps && kkt[1] <= eps && kkt[2] <= eps && cert) s->result = SK_RESULT_OPTIMAL;
    else if (o->time_limit > 0.0 && s->solve_seconds >= o->time_limit) s->result = SK_RESULT_TIME_LIMIT;
    else s->result = SK_RESULT_ITERATION_LIMIT;
    s->dual_bound = s->objective;
    if (stats) *stats = acc;
    return SK_OK;
}

int sk_gpu_gather_rate(int table_elems, int count, int reps, double *gathers_per_second)
{
    int rc;
    skg_clear_error();
    if (table_elems <= 0 || count <= 0 || count > (1 << 30) || reps <= 0 || !gathers_per_second)
        return skg_fail("invalid gather probe arguments");
    if (skg_have_device()) return -1;
    rc = skk_gather_probe(table_elems, count, reps, gathers_per_second);
    return rc ? skg_hip_fail("gather probe", rc) : 0;
}
