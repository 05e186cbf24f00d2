/* CPU oracle -- TEST INFRASTRUCTURE ONLY (see sk_oracle.h).
 *
 * Every routine is an FP64 restatement of a reference routine, written from
 * its behaviour; the citation sits beside each one.
 *
 * How it is pinned (tests/test_oracle.py):
 *   - orc_solve_first_order and orc_verify are compared BIT FOR BIT with the
 *     reference C core compiled from its own sources (oracle/Makefile ->
 *     oracle/_ref) on seeded LPs and QPs, and against the committed golden
 *     fixtures written by tests/golden/make_golden.py;
 *   - orc_verify is also checked on hand-built certificates whose residuals
 *     are known in closed form;
 *   - orc_first_order / orc_pdlp_solve are pinned by the hand-computed optima
 *     of tests/lp/test_pdlp.cpp (x=(2,6) obj -36; x+y=5; row range -> 7;
 *     bounds -> (3,4)) and by the reference simplex optimum of the seeded
 *     bounded LPs (objective within 1e-6 relative);
 *   - orc_cuda_pdhg by cuda/tests/pdhg_smoke.cu and diagonal_qp_smoke.cu
 *     known answers (x=1; x=2, obj -4) and their rejection cases;
 *   - orc_ruiz / orc_pock_chambolle by the invariants of
 *     tests/lp/test_scaling.cpp (row/column inf-norms within tolerance,
 *     zero matrix -> identity) and the closed form on 2x2 matrices;
 *   - orc_power_norm by diagonal and rank-one matrices (closed form) and a
 *     dense SVD on small seeds;
 *   - orc_pdlp_solve_qp / orc_first_order_qp by: an all-zero Hessian, which
 *     must give the LP path's iterates bit for bit (every Q term vanishes);
 *     closed forms (a separable box QP, x = clamp(-c/d); an equality-
 *     constrained two-variable QP solved by its Lagrange system); and the
 *     reference C core's own first-order QP solve (tests/golden/qp_small.npz,
 *     objective within 1e-7 relative, and the same point within the
 *     solve tolerance).
 */
#include "sk_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double wall_seconds(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1.0e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------ CSR */

int orc_csr_from_csc(const sk_csc *a, orc_csr *out)
{
    int j, p, nnz = a->p[a->ncol];
    int *next;
    memset(out, 0, sizeof(*out));
    out->rows = a->nrow; out->cols = a->ncol; out->nnz = nnz;
    out->p = (int *)calloc((size_t)a->nrow + 1, sizeof(int));
    out->j = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    out->v = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    next = (int *)malloc((size_t)(a->nrow + 1) * sizeof(int));
    if (!out->p || !out->j || !out->v || !next) { free(next); orc_csr_free(out); return -1; }
    for (p = 0; p < nnz; ++p) out->p[a->i[p] + 1]++;
    for (j = 0; j < a->nrow; ++j) out->p[j + 1] += out->p[j];
    memcpy(next, out->p, (size_t)(a->nrow + 1) * sizeof(int));
    for (j = 0; j < a->ncol; ++j)
        for (p = a->p[j]; p < a->p[j + 1]; ++p) {
            const int q = next[a->i[p]]++;
            out->j[q] = j; out->v[q] = a->x[p];
        }
    free(next);
    return 0;
}

int orc_csr_transpose(const orc_csr *a, orc_csr *out)
{
    sk_csc as_csc;        /* CSR(A) read as CSC(A^T) */
    as_csc.nrow = a->cols; as_csc.ncol = a->rows;
    as_csc.p = a->p; as_csc.i = a->j; as_csc.x = a->v; as_csc.nzmax = a->nnz;
    return orc_csr_from_csc(&as_csc, out);
}

void orc_csr_free(orc_csr *a)
{
    if (!a) return;
    free(a->p); free(a->j); free(a->v);
    memset(a, 0, sizeof(*a));
}

void orc_csr_mv(const orc_csr *a, const double *x, double *y)
{
    int i, p;
    for (i = 0; i < a->rows; ++i) {
        double s = 0.0;
        for (p = a->p[i]; p < a->p[i + 1]; ++p) s += a->v[p] * x[a->j[p]];
        y[i] = s;
    }
}

/* ------------------------------------------------- model / verification */

void orc_options_default(sk_options *o)      /* c/src/sk_core.c:184-207 */
{
    memset(o, 0, sizeof(*o));
    o->primal_tol = 1e-7; o->dual_tol = 1e-7; o->pivot_tol = 0.1;
    o->presolve = 1; o->scaling = 1; o->refactor_interval = 50;
    o->verbosity = 1; o->threads = 1; o->mip_gap_abs = 1e-9; o->mip_gap_rel = 1e-6;
    o->mip_cuts = 1; o->mip_heuristics = 1; o->dual_simplex = 1;
    o->lp_engine = SK_LP_AUTO; o->random_seed = 20260825ULL;
}

static void solution_init(sk_solution *s)     /* c/src/sk_core.c:166-173 */
{
    memset(s, 0, sizeof(*s));
    s->result = SK_RESULT_NOT_RUN;
    s->objective = NAN;
    s->dual_bound = -SK_INFINITY;
    s->mip_gap = INFINITY;
}

void orc_solution_free(sk_solution *s)
{
    if (!s) return;
    free(s->x); free(s->y); free(s->rc); free(s->rowact);
    solution_init(s);
}

static sk_status csc_validate(const sk_csc *a, int rows, int cols)   /* sk_core.c:126-138 */
{
    int j, p, nnz;
    if (!a || a->nrow != rows || a->ncol != cols || a->nzmax < 0 || !a->p) return SK_ERR_STRUCTURE;
    if (a->p[0] != 0) return SK_ERR_STRUCTURE;
    nnz = a->p[cols];
    if (nnz < 0 || nnz > a->nzmax || (nnz > 0 && (!a->i || !a->x))) return SK_ERR_STRUCTURE;
    for (j = 0; j < cols; ++j)
        if (a->p[j] < 0 || a->p[j] > a->p[j + 1] || a->p[j + 1] > nnz) return SK_ERR_STRUCTURE;
    for (p = 0; p < nnz; ++p)
        if (a->i[p] < 0 || a->i[p] >= rows || !isfinite(a->x[p])) return SK_ERR_STRUCTURE;
    return SK_OK;
}

sk_status orc_model_validate(const sk_model *m)     /* sk_core.c:140-155 */
{
    int i;
    if (!m || m->nrow < 0 || m->ncol < 0) return SK_ERR_ARG;
    if (csc_validate(&m->A, m->nrow, m->ncol) != SK_OK) return SK_ERR_STRUCTURE;
    if ((m->ncol && (!m->c || !m->clow || !m->cupp || !m->vartype)) ||
        (m->nrow && (!m->rlow || !m->rupp))) return SK_ERR_STRUCTURE;
    for (i = 0; i < m->ncol; ++i)
        if (!isfinite(m->c[i]) || isnan(m->clow[i]) || isnan(m->cupp[i]) ||
            m->clow[i] > m->cupp[i] ||
            (m->vartype[i] != SK_CONTINUOUS && m->vartype[i] != SK_INTEGER)) return SK_ERR_STRUCTURE;
    for (i = 0; i < m->nrow; ++i)
        if (isnan(m->rlow[i]) || isnan(m->rupp[i]) || m->rlow[i] > m->rupp[i]) return SK_ERR_STRUCTURE;
    if (m->Q && csc_validate(m->Q, m->ncol, m->ncol) != SK_OK) return SK_ERR_STRUCTURE;
    return SK_OK;
}

/* sk_verify, c/src/sk_core.c:394-510.  rc = c + Qx + A'y; a row at its upper
 * activity bound wants y >= 0, at its lower bound y <= 0; variables at
 * their lower bound want rc >= 0.  "At a bound" means within 1e-9. */
sk_status orc_verify(const sk_model *m, sk_solution *s)
{
    int i, j, p;
    double *act, *rc, *qx = NULL;
    double pinf = 0.0, dinf = 0.0, comp = 0.0, pobj, dobj;
    const double bt = 1e-9;
    if (!m || !s || !s->x) return SK_ERR_ARG;
    if (orc_model_validate(m) != SK_OK) return SK_ERR_STRUCTURE;
    act = (double *)calloc((size_t)m->nrow + 1, sizeof(double));
    rc = (double *)calloc((size_t)m->ncol + 1, sizeof(double));
    if (m->Q) qx = (double *)calloc((size_t)m->ncol + 1, sizeof(double));
    if (!act || !rc || (m->Q && !qx)) { free(act); free(rc); free(qx); return SK_ERR_MEMORY; }

    for (j = 0; j < m->ncol; j++)
        for (p = m->A.p[j]; p < m->A.p[j + 1]; p++) act[m->A.i[p]] += m->A.x[p] * s->x[j];
    pobj = m->objshift;
    for (j = 0; j < m->ncol; j++) pobj += m->c[j] * s->x[j];
    if (m->Q) {
        for (j = 0; j < m->Q->ncol; j++)
            for (p = m->Q->p[j]; p < m->Q->p[j + 1]; p++) qx[m->Q->i[p]] += m->Q->x[p] * s->x[j];
        for (j = 0; j < m->ncol; j++) pobj += 0.5 * s->x[j] * qx[j];
    }
    for (i = 0; i < m->nrow; i++) {
        if (!SK_IS_NEG_INF(m->rlow[i]) && m->rlow[i] - act[i] > pinf) pinf = m->rlow[i] - act[i];
        if (!SK_IS_INF(m->rupp[i]) && act[i] - m->rupp[i] > pinf) pinf = act[i] - m->rupp[i];
    }
    for (j = 0; j < m->ncol; j++) {
        if (!isfinite(s->x[j])) { pinf = INFINITY; break; }
        if (!SK_IS_NEG_INF(m->clow[j]) && m->clow[j] - s->x[j] > pinf) pinf = m->clow[j] - s->x[j];
        if (!SK_IS_INF(m->cupp[j]) && s->x[j] - m->cupp[j] > pinf) pinf = s->x[j] - m->cupp[j];
    }
    dobj = m->objshift;
    if (s->y) {
        for (j = 0; j < m->ncol; j++) {
            double r = m->c[j] + (m->Q ? qx[j] : 0.0);
            for (p = m->A.p[j]; p < m->A.p[j + 1]; p++) r += m->A.x[p] * s->y[m->A.i[p]];
            rc[j] = r;
        }
        for (j = 0; j < m->ncol; j++) {
            const int atlo = !SK_IS_NEG_INF(m->clow[j]) && fabs(s->x[j] - m->clow[j]) <= bt;
            const int athi = !SK_IS_INF(m->cupp[j]) && fabs(s->x[j] - m->cupp[j]) <= bt;
            const double r = rc[j];
            if (atlo && athi) {
            } else if (atlo) { if (-r > dinf) dinf = -r; }
            else if (athi) { if (r > dinf) dinf = r; }
            else { if (fabs(r) > dinf) dinf = fabs(r); }
            if (r > 0.0) {
                if (SK_IS_NEG_INF(m->clow[j])) { if (r > dinf) dinf = r; }
                else dobj += r * m->clow[j];
            } else if (r < 0.0) {
                if (SK_IS_INF(m->cupp[j])) { if (-r > dinf) dinf = -r; }
                else dobj += r * m->cupp[j];
            }
            if (r > 0.0 && !SK_IS_NEG_INF(m->clow[j])) {
                const double product = r * fabs(s->x[j] - m->clow[j]);
                if (product > comp) comp = product;
            } else if (r < 0.0 && !SK_IS_INF(m->cupp[j])) {
                const double product = -r * fabs(m->cupp[j] - s->x[j]);
                if (product > comp) comp = product;
            }
        }
        for (i = 0; i < m->nrow; i++) {
            const int atlo = !SK_IS_NEG_INF(m->rlow[i]) && fabs(act[i] - m->rlow[i]) <= bt;
            const int athi = !SK_IS_INF(m->rupp[i]) && fabs(act[i] - m->rupp[i]) <= bt;
            const double yv = s->y[i];
            if (atlo && athi) {
            } else if (atlo) { if (yv > dinf) dinf = yv; }
            else if (athi) { if (-yv > dinf) dinf = -yv; }
            else { if (fabs(yv) > dinf) dinf = fabs(yv); }
            if (yv < 0.0) {
                if (SK_IS_NEG_INF(m->rlow[i])) { if (-yv > dinf) dinf = -yv; }
                else dobj -= yv * m->rlow[i];
                if (!SK_IS_NEG_INF(m->rlow[i])) {
                    const double product = -yv * fabs(act[i] - m->rlow[i]);
                    if (product > comp) comp = product;
                }
            } else if (yv > 0.0) {
                if (SK_IS_INF(m->rupp[i])) { if (yv > dinf) dinf = yv; }
                else dobj -= yv * m->rupp[i];
                if (!SK_IS_INF(m->rupp[i])) {
                    const double product = yv * fabs(m->rupp[i] - act[i]);
                    if (product > comp) comp = product;
                }
            }
        }
        s->complementarity = m->Q ? comp / (1.0 + fabs(pobj)) : fabs(pobj - dobj) / (1.0 + fabs(pobj));
    } else {
        dinf = INFINITY;
        s->complementarity = INFINITY;
    }
    if (s->rowact) for (i = 0; i < m->nrow; i++) s->rowact[i] = act[i];
    if (s->rc) for (j = 0; j < m->ncol; j++) s->rc[j] = rc[j];
    s->objective = pobj;
    s->primal_infeasibility = pinf > 0.0 ? pinf : 0.0;
    s->dual_infeasibility = dinf > 0.0 ? dinf : 0.0;
    free(act); free(rc); free(qx);
    return SK_OK;
}

/* --------------------------------------------------------- CPU PDHG path */

static double clamp_value(double x, double lo, double hi)    /* sk_solve.c:19-24 */
{
    if (!SK_IS_NEG_INF(lo) && x < lo) x = lo;
    if (!SK_IS_INF(hi) && x > hi) x = hi;
    return x;
}

static void csc_mv(const sk_csc *a, const double *x, double *y)   /* sk_solve.c:26-32 */
{
    int j, p;
    memset(y, 0, (size_t)a->nrow * sizeof(double));
    for (j = 0; j < a->ncol; ++j)
        for (p = a->p[j]; p < a->p[j + 1]; ++p) y[a->i[p]] += a->x[p] * x[j];
}

static void csc_tmv(const sk_csc *a, const double *y, double *x)  /* sk_solve.c:34-58, serial */
{
    int j, p;
    for (j = 0; j < a->ncol; ++j) {
        double sum = 0.0;
        for (p = a->p[j]; p < a->p[j + 1]; ++p) sum += a->x[p] * y[a->i[p]];
        x[j] = sum;
    }
}

static double matrix_norm_bound(const sk_csc *a)   /* sk_solve.c:60-80 */
{
    int j, p;
    double *row_sum, col_max = 0.0, row_max = 0.0;
    if (!a || !a->nrow || !a->ncol) return 0.0;
    row_sum = (double *)calloc((size_t)a->nrow, sizeof(double));
    if (!row_sum) return INFINITY;
    for (j = 0; j < a->ncol; ++j) {
        double col_sum = 0.0;
        for (p = a->p[j]; p < a->p[j + 1]; ++p) {
            const double value = fabs(a->x[p]);
            col_sum += value;
            row_sum[a->i[p]] += value;
        }
        if (col_sum > col_max) col_max = col_sum;
    }
    for (j = 0; j < a->nrow; ++j) if (row_sum[j] > row_max) row_max = row_sum[j];
    free(row_sum);
    return sqrt(col_max * row_max);
}

static double row_violation(const sk_model *m, const double *activity)   /* sk_solve.c:82-93 */
{
    int i;
    double violation = 0.0;
    for (i = 0; i < m->nrow; ++i) {
        if (!SK_IS_NEG_INF(m->rlow[i]) && m->rlow[i] - activity[i] > violation)
            violation = m->rlow[i] - activity[i];
        if (!SK_IS_INF(m->rupp[i]) && activity[i] - m->rupp[i] > violation)
            violation = activity[i] - m->rupp[i];
    }
    return violation;
}

int orc_interval_infeasible(const sk_model *m, double tolerance)   /* sk_solve.c:97-132 */
{
    int i, j, p, impossible = 0;
    double *mn, *mx;
    unsigned char *mn_inf, *mx_inf;
    for (j = 0; j < m->ncol; ++j)
        if (m->clow[j] > m->cupp[j] + tolerance) return 1;
    mn = (double *)calloc((size_t)m->nrow + 1, sizeof(double));
    mx = (double *)calloc((size_t)m->nrow + 1, sizeof(double));
    mn_inf = (unsigned char *)calloc((size_t)m->nrow + 1, 1);
    mx_inf = (unsigned char *)calloc((size_t)m->nrow + 1, 1);
    if (!mn || !mx || !mn_inf || !mx_inf) { free(mn); free(mx); free(mn_inf); free(mx_inf); return 0; }
    for (j = 0; j < m->ncol; ++j)
        for (p = m->A.p[j]; p < m->A.p[j + 1]; ++p) {
            const int row = m->A.i[p];
            const double value = m->A.x[p];
            const double lo = value >= 0.0 ? m->clow[j] : m->cupp[j];
            const double hi = value >= 0.0 ? m->cupp[j] : m->clow[j];
            if ((value >= 0.0 && SK_IS_NEG_INF(lo)) || (value < 0.0 && SK_IS_INF(lo))) mn_inf[row] = 1;
            else if (!mn_inf[row]) mn[row] += value * lo;
            if ((value >= 0.0 && SK_IS_INF(hi)) || (value < 0.0 && SK_IS_NEG_INF(hi))) mx_inf[row] = 1;
            else if (!mx_inf[row]) mx[row] += value * hi;
        }
    for (i = 0; i < m->nrow; ++i) {
        if (!mx_inf[i] && m->rlow[i] > mx[i] + tolerance) { impossible = 1; break; }
        if (!mn_inf[i] && m->rupp[i] < mn[i] - tolerance) { impossible = 1; break; }
    }
    free(mn); free(mx); free(mn_inf); free(mx_inf);
    return impossible;
}

/* solve_continuous (sk_solve.c:801-1007) on the models where none of the
 * small-QP shortcuts apply: LPs, and QPs with n > 512 and a nonempty A
 * (the PSD check, closed forms, enumeration, polish and working set are all
 * gated below those sizes, sk_solve.c:206, 379, 474, 608, 834).  QPs that
 * would take a shortcut are reported SK_ERR_UNSUPPORTED here. */
static sk_status solve_continuous(const sk_model *m, const sk_options *o, sk_solution *s)
{
    const int n = m->ncol, r = m->nrow;
    const int maximum_iterations = (o->iteration_limit > 0 && o->iteration_limit < 2147483647LL)
        ? (int)o->iteration_limit : 100000;
    const int check_every = 20;
    double *x, *x_new, *x_bar, *y, *y_new, *activity, *gradient, *qx = NULL, *qdiag = NULL;
    double *tau_vec, *sigma_vec, anorm, qnorm, start, dual_step = INFINITY;
    int iteration, performed_iterations, converged = 0, j, p;

    if (m->Q && (m->Q->nrow != n || m->Q->ncol != n)) return SK_ERR_UNSUPPORTED;
    if (o->primal_tol <= 0.0 || o->dual_tol <= 0.0) return SK_ERR_ARG;
    if (orc_interval_infeasible(m, o->primal_tol)) {
        solution_init(s);
        s->result = SK_RESULT_INFEASIBLE; s->ncol = n; s->nrow = r;
        s->primal_infeasibility = INFINITY;
        return SK_OK;
    }
    if (m->Q && (n <= 512 || m->A.p[n] == 0 || n + r <= 256)) return SK_ERR_UNSUPPORTED;

    x = (double *)calloc((size_t)n + 1, sizeof(double));
    x_new = (double *)calloc((size_t)n + 1, sizeof(double));
    x_bar = (double *)calloc((size_t)n + 1, sizeof(double));
    y = (double *)calloc((size_t)r + 1, sizeof(double));
    y_new = (double *)calloc((size_t)r + 1, sizeof(double));
    activity = (double *)calloc((size_t)r + 1, sizeof(double));
    gradient = (double *)calloc((size_t)n + 1, sizeof(double));
    tau_vec = (double *)calloc((size_t)n + 1, sizeof(double));
    sigma_vec = (double *)calloc((size_t)r + 1, sizeof(double));
    if (m->Q) { qx = (double *)calloc((size_t)n + 1, sizeof(double)); qdiag = (double *)calloc((size_t)n + 1, sizeof(double)); }
    if (!x || !x_new || !x_bar || !y || !y_new || !activity || !gradient || !tau_vec || !sigma_vec ||
        (m->Q && (!qx || !qdiag))) goto memory_failure;
    {
        int diagonal = m->Q != NULL;
        if (m->Q) for (j = 0; j < n; ++j) for (p = m->Q->p[j]; p < m->Q->p[j + 1]; ++p) {
            if (m->Q->i[p] != j) { diagonal = 0; break; }
            qdiag[j] += m->Q->x[p];
        }
        if (!diagonal) { free(qdiag); qdiag = NULL; }
    }
    for (j = 0; j < n; ++j) x[j] = x_bar[j] = clamp_value(0.0, m->clow[j], m->cupp[j]);
    anorm = matrix_norm_bound(&m->A);
    qnorm = matrix_norm_bound(m->Q);
    if (!isfinite(anorm) || !isfinite(qnorm)) goto memory_failure;
    {
        double *row_mass = (double *)calloc((size_t)r + 1, sizeof(double));
        double *col_mass = (double *)calloc((size_t)n + 1, sizeof(double));
        if (!row_mass || !col_mass) { free(row_mass); free(col_mass); goto memory_failure; }
        for (j = 0; j < n; ++j) {
            for (p = m->A.p[j]; p < m->A.p[j + 1]; ++p) {
                row_mass[m->A.i[p]] += fabs(m->A.x[p]);
                col_mass[j] += fabs(m->A.x[p]);
            }
            if (m->Q) for (p = m->Q->p[j]; p < m->Q->p[j + 1]; ++p) col_mass[j] += fabs(m->Q->x[p]);
        }
        for (j = 0; j < n; ++j) tau_vec[j] = 0.9 / (1.0 + col_mass[j]);
        for (j = 0; j < r; ++j) sigma_vec[j] = 0.9 / (1.0 + row_mass[j]);
        free(row_mass); free(col_mass);
    }
    start = wall_seconds();
    for (iteration = 1; iteration <= maximum_iterations; ++iteration) {
        csc_mv(&m->A, x_bar, activity);
        for (j = 0; j < r; ++j) {
            const double trial = y[j] + sigma_vec[j] * activity[j];
            y_new[j] = trial - sigma_vec[j] * clamp_value(trial / sigma_vec[j], m->rlow[j], m->rupp[j]);
        }
        dual_step = 0.0;
        for (j = 0; j < r; ++j) if (fabs(y_new[j] - y[j]) > dual_step) dual_step = fabs(y_new[j] - y[j]);
        csc_tmv(&m->A, y_new, gradient);
        if (m->Q) {
            if (qdiag) for (j = 0; j < n; ++j) qx[j] = qdiag[j] * x[j];
            else csc_mv(m->Q, x, qx);
        }
        for (j = 0; j < n; ++j) {
            if (qdiag) x_new[j] = (x[j] - tau_vec[j] * (m->c[j] + gradient[j])) / (1.0 + tau_vec[j] * qdiag[j]);
            else x_new[j] = x[j] - tau_vec[j] * (m->c[j] + (m->Q ? qx[j] : 0.0) + gradient[j]);
            x_new[j] = clamp_value(x_new[j], m->clow[j], m->cupp[j]);
            x_bar[j] = 2.0 * x_new[j] - x[j];
        }
        memcpy(x, x_new, (size_t)n * sizeof(double));
        memcpy(y, y_new, (size_t)r * sizeof(double));
        if (o->time_limit > 0.0 && wall_seconds() - start >= o->time_limit) break;
        if (iteration % check_every == 0 || iteration == maximum_iterations) {
            double step = 0.0, scale = 1.0;
            int qp_kkt_ready = 1;
            csc_mv(&m->A, x, activity);
            for (j = 0; j < n; ++j) if (fabs(x[j]) > scale) scale = fabs(x[j]);
            for (j = 0; j < n; ++j) if (fabs(x_bar[j] - x[j]) > step) step = fabs(x_bar[j] - x[j]);
            if (m->Q) {
                sk_solution probe;
                double score;
                solution_init(&probe);
                probe.x = x; probe.y = y;
                score = orc_verify(m, &probe) == SK_OK
                    ? probe.primal_infeasibility + probe.dual_infeasibility + probe.complementarity : INFINITY;
                qp_kkt_ready = isfinite(score) && score <= 100.0 * (o->primal_tol + 2.0 * o->dual_tol);
            }
            if (row_violation(m, activity) <= o->primal_tol && step <= o->primal_tol * scale &&
                dual_step <= o->dual_tol * (1.0 + anorm) && qp_kkt_ready) {
                converged = 1;
                break;
            }
        }
    }
    performed_iterations = iteration > maximum_iterations ? maximum_iterations : iteration;
    solution_init(s);
    s->x = x; x = NULL;
    s->y = y; y = NULL;
    s->rowact = activity; activity = NULL;
    s->ncol = n; s->nrow = r;
    s->iterations = performed_iterations;
    s->solve_seconds = wall_seconds() - start;
    if (m->Q) {
        if (qdiag) for (j = 0; j < n; ++j) qx[j] = qdiag[j] * s->x[j];
        else csc_mv(m->Q, s->x, qx);
    }
    s->objective = m->objshift;
    for (j = 0; j < n; ++j)
        s->objective += m->c[j] * s->x[j] + (m->Q ? 0.5 * s->x[j] * qx[j] : 0.0);
    s->primal_infeasibility = row_violation(m, s->rowact);
    if (!isfinite(s->objective) || !isfinite(s->primal_infeasibility)) {
        s->result = SK_RESULT_NUMERIC_FAILURE;
    } else if (m->Q) {
        if (orc_verify(m, s) == SK_OK && s->primal_infeasibility <= 100.0 * o->primal_tol &&
            s->dual_infeasibility <= 100.0 * o->dual_tol && s->complementarity <= 100.0 * o->dual_tol)
            converged = 1;
    }
    if (s->result != SK_RESULT_NUMERIC_FAILURE && converged &&
        (!m->Q || (s->primal_infeasibility <= 100.0 * o->primal_tol &&
                   s->dual_infeasibility <= 100.0 * o->dual_tol && s->complementarity <= 100.0 * o->dual_tol))) {
        s->result = SK_RESULT_OPTIMAL;
    } else {
        s->result = o->time_limit > 0.0 && s->solve_seconds >= o->time_limit
            ? SK_RESULT_TIME_LIMIT : SK_RESULT_ITERATION_LIMIT;
    }
    free(x_new); free(x_bar); free(y_new); free(gradient); free(qx); free(qdiag); free(tau_vec); free(sigma_vec);
    return SK_OK;
memory_failure:
    free(x); free(x_new); free(x_bar); free(y); free(y_new); free(activity); free(gradient); free(qx);
    free(qdiag); free(tau_vec); free(sigma_vec);
    return SK_ERR_MEMORY;
}

sk_status orc_solve_first_order(const sk_model *m, const sk_options *options, sk_solution *s)
{                                                        /* sk_solve.c:1009-1046 */
    sk_options fallback;
    int j;
    if (!m || !s) return SK_ERR_ARG;
    if (orc_model_validate(m) != SK_OK) return SK_ERR_STRUCTURE;
    if (!options) { orc_options_default(&fallback); options = &fallback; }
    if (options->lp_engine < SK_LP_AUTO || options->lp_engine > SK_LP_FIRST_ORDER) return SK_ERR_ARG;
    if (orc_interval_infeasible(m, options->primal_tol > 0.0 ? options->primal_tol : 1e-7)) {
        solution_init(s);
        s->result = SK_RESULT_INFEASIBLE; s->ncol = m->ncol; s->nrow = m->nrow;
        s->primal_infeasibility = INFINITY;
        return SK_OK;
    }
    for (j = 0; j < m->ncol; ++j) if (m->vartype[j] == SK_INTEGER) return SK_ERR_UNSUPPORTED;
    return solve_continuous(m, options, s);
}

/* ------------------------------------- plain PDHG with host checkpoints */

static double project_box(double value, double lower, double upper)   /* sankhya_cuda_lp.cu:13-17 */
{
    if (!isinf(lower) && value < lower) value = lower;
    if (!isinf(upper) && value > upper) value = upper;
    return value;
}

static int valid_lower(double v) { return !isnan(v) && !(isinf(v) && v > 0.0); }
static int valid_upper(double v) { return !isnan(v) && !(isinf(v) && v < 0.0); }
static int has_lower(double v) { return isfinite(v) && v > -5.0e29; }
static int has_upper(double v) { return isfinite(v) && v < 5.0e29; }

static void csr_mv_fma(const orc_csr *a, const double *x, double *y)  /* sankhya_cuda.cu:28-38 */
{
    int i, p;
    for (i = 0; i < a->rows; ++i) {
        double sum = 0.0;
        for (p = a->p[i]; p < a->p[i + 1]; ++p) sum = fma(a->v[p], x[a->j[p]], sum);
        y[i] = sum;
    }
}

/* cuda/src/sankhya_cuda_lp.cu:101-385.  The transpose product is a gather
 * over CSR(A^T) in ascending row order, one of the orders the reference's
 * atomic scatter (sankhya_cuda.cu:45-53) can produce. */
int orc_cuda_pdhg(const orc_csr *A, const orc_csr *Q, const double *qd,
                  const double *c, const double *rl, const double *ru,
                  const double *cl, const double *cu, sk_gpu_lp_settings st,
                  const double *ps, const double *ds, double *solution, double *dual_solution,
                  sk_gpu_lp_result *res)
{
    int rows, cols, i, j, it, final_iteration, have_prev = 0, rc = 0;
    orc_csr At;
    double *x, *xn, *xb, *y, *yn, *act, *grad, *qg = NULL, *prev;
    double start;
    if (!res) return -1;
    res->status = -1; res->iterations = 0; res->objective = NAN;
    res->maximum_row_violation = res->maximum_step = res->maximum_kkt_residual = INFINITY;
    res->maximum_dual_residual = res->maximum_complementarity = INFINITY;
    if (!A || !c || !rl || !ru || !cl || !cu || !solution) return -1;
    if (st.max_iterations <= 0 || st.check_every <= 0 || st.theta < 0.0 || st.theta > 1.0 ||
        st.tolerance <= 0.0 || st.time_limit < 0.0) return -1;
    if ((ps == NULL) != (ds == NULL)) return -1;
    if (ps == NULL && (st.tau <= 0.0 || st.sigma <= 0.0)) return -1;
    rows = A->rows; cols = A->cols;
    for (j = 0; j < cols; ++j)
        if (!isfinite(c[j]) || !valid_lower(cl[j]) || !valid_upper(cu[j]) || cl[j] > cu[j]) return -1;
    for (i = 0; i < rows; ++i)
        if (!valid_lower(rl[i]) || !valid_upper(ru[i]) || rl[i] > ru[i]) return -1;
    if (ps) {
        for (j = 0; j < cols; ++j) if (!isfinite(ps[j]) || ps[j] <= 0.0) return -1;
        for (i = 0; i < rows; ++i) if (!isfinite(ds[i]) || ds[i] <= 0.0) return -1;
    }
    if (Q && (qd || Q->rows != cols || Q->cols != cols)) return -1;
    if (qd) for (j = 0; j < cols; ++j) if (!isfinite(qd[j]) || qd[j] < 0.0) return -1;
    if (orc_csr_transpose(A, &At)) return -1;
    x = (double *)calloc((size_t)cols + 1, sizeof(double));
    xn = (double *)calloc((size_t)cols + 1, sizeof(double));
    xb = (double *)calloc((size_t)cols + 1, sizeof(double));
    grad = (double *)calloc((size_t)cols + 1, sizeof(double));
    prev = (double *)calloc((size_t)cols + 1, sizeof(double));
    y = (double *)calloc((size_t)rows + 1, sizeof(double));
    yn = (double *)calloc((size_t)rows + 1, sizeof(double));
    act = (double *)calloc((size_t)rows + 1, sizeof(double));
    if (Q) qg = (double *)calloc((size_t)cols + 1, sizeof(double));
    for (j = 0; j < cols; ++j) x[j] = xb[j] = project_box(0.0, cl[j], cu[j]);
    final_iteration = st.max_iterations;
    start = wall_seconds();
    for (it = 1; it <= st.max_iterations; ++it) {
        if (st.time_limit > 0.0 && it % st.check_every == 1 && wall_seconds() - start >= st.time_limit) {
            final_iteration = it - 1; res->status = 2; break;
        }
        csr_mv_fma(A, xb, act);
        for (i = 0; i < rows; ++i) {
            const double sigma = ds ? ds[i] : st.sigma;
            const double trial = y[i] + sigma * act[i];
            yn[i] = trial - sigma * project_box(trial / sigma, rl[i], ru[i]);
        }
        for (j = 0; j < cols; ++j) {              /* x_j = sum_i a_ij * y_i, rounded products */
            double s = 0.0;
            int p;
            for (p = At.p[j]; p < At.p[j + 1]; ++p) s += At.v[p] * yn[At.j[p]];
            grad[j] = s;
        }
        if (Q) csr_mv_fma(Q, x, qg);
        for (j = 0; j < cols; ++j) {
            const double tau = ps ? ps[j] : st.tau;
            const double curv = qd ? qd[j] : 0.0;
            const double qgr = qg ? qg[j] : 0.0;
            const double trial = (x[j] - tau * (c[j] + grad[j] + qgr)) / (1.0 + tau * curv);
            const double next = project_box(trial, cl[j], cu[j]);
            xn[j] = next;
            xb[j] = next + st.theta * (next - x[j]);
        }
        { double *t = x; x = xn; xn = t; t = y; y = yn; yn = t; }
        if (it % st.check_every == 0 || it == st.max_iterations) {
            double objective = 0.0, step, scale = 1.0, viol = 0.0, kkt = 0.0, dres = 0.0, comp = 0.0;
            csr_mv_fma(A, x, act);
            for (j = 0; j < cols; ++j) {
                double s = 0.0;
                int p;
                for (p = At.p[j]; p < At.p[j + 1]; ++p) s += At.v[p] * y[At.j[p]];
                grad[j] = s;
            }
            for (j = 0; j < cols; ++j) if (!isfinite(x[j]) || !isfinite(grad[j])) { rc = -2; break; }
            for (i = 0; i < rows && !rc; ++i) if (!isfinite(act[i]) || !isfinite(y[i])) { rc = -2; break; }
            if (!rc && Q) {
                csr_mv_fma(Q, x, qg);
                for (j = 0; j < cols; ++j) if (!isfinite(qg[j])) { rc = -2; break; }
            }
            if (!rc) {
                for (j = 0; j < cols; ++j) {
                    objective += c[j] * x[j];
                    if (qd) objective += 0.5 * qd[j] * x[j] * x[j];
                    else if (Q) objective += 0.5 * x[j] * qg[j];
                }
                if (!isfinite(objective)) rc = -2;
            }
            if (rc) { res->status = -2; break; }
            step = have_prev ? 0.0 : INFINITY;
            for (j = 0; j < cols; ++j) {
                if (have_prev && fabs(x[j] - prev[j]) > step) step = fabs(x[j] - prev[j]);
                if (fabs(x[j]) > scale) scale = fabs(x[j]);
                prev[j] = x[j];
            }
            have_prev = 1;
            for (i = 0; i < rows; ++i) {
                if (isfinite(rl[i]) && rl[i] - act[i] > viol) viol = rl[i] - act[i];
                if (isfinite(ru[i]) && act[i] - ru[i] > viol) viol = act[i] - ru[i];
            }
            for (i = 0; i < rows; ++i) {
                const double projected = project_box(act[i] + y[i], rl[i], ru[i]);
                if (fabs(act[i] - projected) > kkt) kkt = fabs(act[i] - projected);
            }
            for (j = 0; j < cols; ++j) {
                const double qgr = qd ? qd[j] * x[j] : (Q ? qg[j] : 0.0);
                const double g = c[j] + grad[j] + qgr;
                const double projected = project_box(x[j] - g, cl[j], cu[j]);
                if (fabs(x[j] - projected) > kkt) kkt = fabs(x[j] - projected);
            }
            for (j = 0; j < cols; ++j) {
                const double qgr = qd ? qd[j] * x[j] : (Q ? qg[j] : 0.0);
                const double r = c[j] + grad[j] + qgr;
                const int atlo = has_lower(cl[j]) && fabs(x[j] - cl[j]) <= 1e-9;
                const int athi = has_upper(cu[j]) && fabs(x[j] - cu[j]) <= 1e-9;
                if (atlo && athi) {
                } else if (atlo) { if (-r > dres) dres = -r; }
                else if (athi) { if (r > dres) dres = r; }
                else { if (fabs(r) > dres) dres = fabs(r); }
                if (r > 0.0 && has_lower(cl[j])) { const double t = r * fabs(x[j] - cl[j]); if (t > comp) comp = t; }
                else if (r < 0.0 && has_upper(cu[j])) { const double t = -r * fabs(cu[j] - x[j]); if (t > comp) comp = t; }
            }
            for (i = 0; i < rows; ++i) {
                const double mu = y[i], a = act[i];
                const int atlo = has_lower(rl[i]) && fabs(a - rl[i]) <= 1e-9;
                const int athi = has_upper(ru[i]) && fabs(a - ru[i]) <= 1e-9;
                if (atlo && athi) {
                } else if (atlo) { if (mu > dres) dres = mu; }
                else if (athi) { if (-mu > dres) dres = -mu; }
                else { if (fabs(mu) > dres) dres = fabs(mu); }
                if (mu < 0.0 && has_lower(rl[i])) { const double t = -mu * fabs(a - rl[i]); if (t > comp) comp = t; }
                else if (mu > 0.0 && has_upper(ru[i])) { const double t = mu * fabs(ru[i] - a); if (t > comp) comp = t; }
            }
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
    if (!rc) {
        if (res->status < 0) res->status = 1;
        res->iterations = final_iteration;
        memcpy(solution, x, (size_t)cols * sizeof(double));
        if (dual_solution) memcpy(dual_solution, y, (size_t)rows * sizeof(double));
    }
    free(x); free(xn); free(xb); free(grad); free(prev); free(y); free(yn); free(act); free(qg);
    orc_csr_free(&At);
    return rc;
}

/* --------------------------------------------------------------- scaling */

void orc_ruiz(const orc_csr *a, int max_iters, double tol, double *R, double *C)
{                                                        /* Scaling.cpp:11-92 */
    const int m = a->rows, n = a->cols;
    int i, j, k, iter, converged = 0;
    orc_csr at;
    double *rn, *cn;
    for (i = 0; i < m; ++i) R[i] = 1.0;
    for (j = 0; j < n; ++j) C[j] = 1.0;
    if (a->nnz == 0 || m == 0 || n == 0) return;
    if (orc_csr_transpose(a, &at)) return;
    rn = (double *)malloc((size_t)m * sizeof(double));
    cn = (double *)malloc((size_t)n * sizeof(double));
    for (iter = 0; iter < max_iters; ++iter) {
        double max_dev = 0.0;
        for (i = 0; i < m; ++i) {
            double best = 0.0;
            for (k = a->p[i]; k < a->p[i + 1]; ++k) {
                const double v = fabs(a->v[k]) * C[a->j[k]];
                if (v > best) best = v;
            }
            rn[i] = R[i] * best;
        }
        for (j = 0; j < n; ++j) {
            double best = 0.0;
            for (k = at.p[j]; k < at.p[j + 1]; ++k) {
                const double v = fabs(at.v[k]) * R[at.j[k]];
                if (v > best) best = v;
            }
            cn[j] = C[j] * best;
        }
        for (i = 0; i < m; ++i) if (rn[i] > 0.0 && fabs(rn[i] - 1.0) > max_dev) max_dev = fabs(rn[i] - 1.0);
        for (j = 0; j < n; ++j) if (cn[j] > 0.0 && fabs(cn[j] - 1.0) > max_dev) max_dev = fabs(cn[j] - 1.0);
        if (max_dev < tol) { converged = 1; break; }
        for (i = 0; i < m; ++i) if (rn[i] > 0.0) R[i] /= sqrt(rn[i]);
        for (j = 0; j < n; ++j) if (cn[j] > 0.0) C[j] /= sqrt(cn[j]);
    }
    if (!converged) {
        for (i = 0; i < m; ++i) R[i] = 1.0;
        for (j = 0; j < n; ++j) C[j] = 1.0;
    }
    free(rn); free(cn);
    orc_csr_free(&at);
}

void orc_pock_chambolle(const orc_csr *a, double alpha, double *R, double *C)
{                                                        /* Scaling.cpp:111-145 */
    int i, j, k;
    double *ra = (double *)calloc((size_t)a->rows + 1, sizeof(double));
    double *ca = (double *)calloc((size_t)a->cols + 1, sizeof(double));
    const double row_exp = 2.0 - alpha;
    for (i = 0; i < a->rows; ++i)
        for (k = a->p[i]; k < a->p[i + 1]; ++k) {
            const double v = fabs(a->v[k]);
            if (v == 0.0) continue;
            ra[i] += pow(v, row_exp);
            ca[a->j[k]] += pow(v, alpha);
        }
    for (i = 0; i < a->rows; ++i) R[i] = ra[i] > 0.0 ? 1.0 / sqrt(ra[i]) : 1.0;
    for (j = 0; j < a->cols; ++j) C[j] = ca[j] > 0.0 ? 1.0 / sqrt(ca[j]) : 1.0;
    free(ra); free(ca);
}

/* ------------------------------------------------------------------ PDLP */

double orc_power_norm(const orc_csr *a, int iterations)   /* GpuPdlp.cpp:179-211 */
{
    const int n = a->cols, m = a->rows;
    int j, k;
    orc_csr at;
    double *v, *w, *v2, s = 0.0, nrm, lambda;
    if (orc_csr_transpose(a, &at)) return 1.0;
    v = (double *)malloc((size_t)(n + 1) * sizeof(double));
    w = (double *)malloc((size_t)(m + 1) * sizeof(double));
    v2 = (double *)malloc((size_t)(n + 1) * sizeof(double));
    for (j = 0; j < n; ++j) v[j] = 1.0;
    for (j = 0; j < n; ++j) s += v[j] * v[j];
    nrm = sqrt(s);
    if (nrm > 0.0) for (j = 0; j < n; ++j) v[j] *= 1.0 / nrm;
    s = 0.0;
    for (k = 0; k < iterations; ++k) {
        orc_csr_mv(a, v, w);
        orc_csr_mv(&at, w, v2);
        s = 0.0;
        for (j = 0; j < n; ++j) s += v2[j] * v2[j];
        memcpy(v, v2, (size_t)n * sizeof(double));
        nrm = sqrt(s);
        if (nrm > 0.0) for (j = 0; j < n; ++j) v[j] *= 1.0 / nrm;
    }
    free(v); free(w); free(v2);
    orc_csr_free(&at);
    lambda = sqrt(s > 0.0 ? s : 0.0);
    nrm = sqrt(lambda);
    return (nrm > 0.0 && isfinite(nrm)) ? nrm : 1.0;
}

typedef struct { double primal, dual, gap, pobj, dobj, pmove, dmove; } kkt_t;

static double kkt_worst(const kkt_t *k)
{
    const double a = k->primal > k->dual ? k->primal : k->dual;
    return a > k->gap ? a : k->gap;
}

static double clamp_to(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

/* PdlpKernels.cu:101-209 (k_residuals) + GpuPdlp.cpp:226-249, sums in
 * ascending index order; column terms first, then row terms. */
static kkt_t residuals(int n, int m, const double *x, const double *y, const double *ax,
                       const double *aty, const double *c, const double *l, const double *u,
                       const double *rl, const double *ru, const double *xr, const double *yr,
                       double cost_norm, double bound_norm)
{
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    kkt_t k;
    int j, i;
    for (j = 0; j < n; ++j) {
        const double xj = x[j], lo = l[j], hi = u[j], r = c[j] + aty[j], dx = xj - xr[j];
        double viol = 0.0, term = 0.0;
        acc[2] += c[j] * xj;
        acc[4] += xj * xj;
        acc[6] += dx * dx;
        if (isfinite(lo) && isfinite(hi)) term = fmin(r * lo, r * hi);
        else if (isfinite(lo)) { viol = fmax(0.0, -r); term = fmax(0.0, r) * lo; }
        else if (isfinite(hi)) { viol = fmax(0.0, r); term = fmin(0.0, r) * hi; }
        else viol = fabs(r);
        acc[1] += viol * viol;
        acc[3] += term;
    }
    for (i = 0; i < m; ++i) {
        const double a = ax[i], pr = a - clamp_to(a, rl[i], ru[i]), yi = y[i], dy = yi - yr[i];
        double support = 0.0;
        acc[0] += pr * pr;
        acc[5] += yi * yi;
        acc[7] += dy * dy;
        if (yi > 0.0) support = isfinite(ru[i]) ? yi * ru[i] : 0.0;
        else if (yi < 0.0) support = isfinite(rl[i]) ? yi * rl[i] : 0.0;
        acc[3] -= support;
    }
    k.primal = sqrt(fmax(0.0, acc[0])) / (1.0 + bound_norm);
    k.dual = sqrt(fmax(0.0, acc[1])) / (1.0 + cost_norm);
    k.pobj = acc[2]; k.dobj = acc[3];
    k.gap = fabs(acc[2] - acc[3]) / (1.0 + fabs(acc[2]) + fabs(acc[3]));
    k.pmove = sqrt(fmax(0.0, acc[6]));
    k.dmove = sqrt(fmax(0.0, acc[7]));
    if (!isfinite(k.primal) || !isfinite(k.dual) || !isfinite(k.gap))
        k.primal = k.dual = k.gap = INFINITY;
    return k;
}

/* residuals() of a QP: the dual residual and the dual terms use c + Qx, and
 * both objectives carry x'Qx/2 (primal c'x + x'Qx/2, dual ... - x'Qx/2);
 * qx = Q x (NULL: an LP, residuals() itself).  `cq` is n doubles of
 * scratch. */
static kkt_t residuals_q(int n, int m, const double *x, const double *y, const double *ax,
                         const double *aty, const double *qx, const double *c, const double *l,
                         const double *u, const double *rl, const double *ru, const double *xr,
                         const double *yr, double cost_norm, double bound_norm, double *cq)
{
    kkt_t k;
    double xqx = 0.0;
    int j;
    if (!qx) return residuals(n, m, x, y, ax, aty, c, l, u, rl, ru, xr, yr, cost_norm, bound_norm);
    for (j = 0; j < n; ++j) { cq[j] = c[j] + qx[j]; xqx += x[j] * qx[j]; }
    k = residuals(n, m, x, y, ax, aty, cq, l, u, rl, ru, xr, yr, cost_norm, bound_norm);
    k.pobj -= 0.5 * xqx;
    k.dobj -= 0.5 * xqx;
    k.gap = fabs(k.pobj - k.dobj) / (1.0 + fabs(k.pobj) + fabs(k.dobj));
    if (!isfinite(k.primal) || !isfinite(k.dual) || !isfinite(k.gap))
        k.primal = k.dual = k.gap = INFINITY;
    return k;
}

static double finite_norm(const double *v, int n)   /* GpuPdlp.cpp:84-90 */
{
    double acc = 0.0;
    int k;
    for (k = 0; k < n; ++k) if (isfinite(v[k])) acc += v[k] * v[k];
    return sqrt(acc);
}

/* ---------------------------------------------- normalized duality gap
 *
 * Restart merit of Applegate et al. 2021 (PDLP, s.3.1), not in the
 * reference (GpuPdlp.cpp:437-451 restarts on the relative KKT).  For the
 * LP saddle function L(x, y) = c'x + y'Ax - p(y), p the support function
 * of [rl, ru], the gap of z = (x, y) within the omega-ball of radius
 * r = ||z - z_r||_omega around z (z_r: the last restart point) is
 *     G = max { L(x, yh) - L(xh, y) : xh in [l, u], ||zh - z||_omega <= r }
 * and the merit is G / r.  With a multiplier on the ball the maximiser is
 * one projected step of length t: xh = proj(x - t g / omega), g = c + A'y,
 * and yh = prox_{s p}(y + s A x) with s = t omega; the distance grows with
 * t, so t is found by bisection on the distance (exact to rounding here;
 * the device uses a three-stage 32-point search).  Written with
 * differences, G = sum_j g_j (x_j - xh_j) + sum_i [(yh_i - y_i)(Ax)_i -
 * p_i(yh_i) + p_i(y_i)], so no objective-sized terms cancel. */
static double support(double v, double lo, double hi)
{
    if (v > 0.0) return isfinite(hi) ? v * hi : 0.0;
    if (v < 0.0) return isfinite(lo) ? v * lo : 0.0;
    return 0.0;
}

static void ndg_at(double t, double omega, int n, int m, const double *x, const double *y, const double *ax,
                   const double *g, const double *l, const double *u, const double *rl, const double *ru,
                   double *d2, double *G)
{
    double dist = 0.0, gap = 0.0;
    int j, i;
    for (j = 0; j < n; ++j) {
        const double d = clamp_to(x[j] - t * g[j] / omega, l[j], u[j]) - x[j];
        dist += omega * d * d;
        gap -= g[j] * d;
    }
    for (i = 0; i < m; ++i) {
        const double s = t * omega, v = y[i] + s * ax[i];
        const double yh = v - s * clamp_to(v / s, rl[i], ru[i]);
        const double d = yh - y[i];
        dist += d * d / omega;
        gap += d * ax[i] - support(yh, rl[i], ru[i]) + support(y[i], rl[i], ru[i]);
    }
    *d2 = dist;
    *G = gap;
}

double orc_ndg(int n, int m, const double *x, const double *y, const double *ax, const double *g,
               const double *l, const double *u, const double *rl, const double *ru, const double *xr,
               const double *yr, double omega)
{
    double r2 = 0.0, r, lo = 0.0, hi, d2, G, gn = 0.0;
    int j, i, k;
    for (j = 0; j < n; ++j) r2 += omega * (x[j] - xr[j]) * (x[j] - xr[j]);
    for (i = 0; i < m; ++i) r2 += (y[i] - yr[i]) * (y[i] - yr[i]) / omega;
    r = sqrt(r2);
    if (!(r > 0.0)) return 0.0;
    for (j = 0; j < n; ++j) gn += g[j] * g[j] / omega;
    for (i = 0; i < m; ++i) gn += omega * ax[i] * ax[i];
    /* d(t) <= t sqrt(gn) (projections are nonexpansive), so t >= r / sqrt(gn) */
    hi = gn > 0.0 ? r / sqrt(gn) : 1.0;
    for (k = 0; k < 200; ++k) {
        ndg_at(hi, omega, n, m, x, y, ax, g, l, u, rl, ru, &d2, &G);
        if (d2 >= r2) break;
        lo = hi;
        hi *= 2.0;
    }
    for (k = 0; k < 100 && lo > 0.0; ++k) {
        const double mid = sqrt(lo * hi);
        ndg_at(mid, omega, n, m, x, y, ax, g, l, u, rl, ru, &d2, &G);
        if (d2 <= r2) lo = mid; else hi = mid;
    }
    if (!(lo > 0.0)) lo = hi;
    ndg_at(lo, omega, n, m, x, y, ax, g, l, u, rl, ru, &d2, &G);
    return G / r;
}

/* GpuPdlp::solve, src/cuda/GpuPdlp.cpp:252-534, with the kernels of
 * src/cuda/PdlpKernels.cu:71-97 (fixed step) and 360-518 (adaptive step). */
static int pdlp_solve(const orc_csr *A, const orc_csr *Q, const double *c, const double *l,
                      const double *u, const double *rl, const double *ru,
                      const sk_gpu_pdlp_params *P, const double *x0, const double *y0,
                      double *x_out, double *y_out, sk_gpu_pdlp_stats *st)
{
    const int n = A->cols, m = A->rows;
    const double t0 = wall_seconds();
    orc_csr At;
    double *x, *xb, *xs, *xr, *xa, *aty, *y, *ys, *yr, *ya, *axb, *ax, *xc, *yc, *axc, *axp;
    double cost_norm, bound_norm, norm, eta_global, eta_min, eta_max, omega = 1.0;
    double eta, accepted_total = 0.0, accepted_at_restart = 0.0, accept_count = 0.0, updates = 0.0;
    double mu_at_restart = INFINITY, mu_previous = INFINITY;
    int iteration = 0, windows_since_restart = 0, its_since_restart = 0, i, j, k;
    double *qx = NULL, *qxc = NULL, *qxa = NULL, *cq = NULL;
    kkt_t best = {1, 1, 1, 0, 0, 0, 0};
    memset(st, 0, sizeof(*st));
    /* the Hessian enters the adaptive-step engine only, and not the NDG
       merit (its trust-region maximiser is not a projection for a QP) */
    if (Q && (!P->adaptive_step || P->restart_rule == SK_GPU_RESTART_NDG)) return -1;
    if (orc_csr_transpose(A, &At)) return -1;
#define ALLOCN(p) p = (double *)calloc((size_t)n + 1, sizeof(double))
#define ALLOCM(p) p = (double *)calloc((size_t)m + 1, sizeof(double))
    ALLOCN(x); ALLOCN(xb); ALLOCN(xs); ALLOCN(xr); ALLOCN(xa); ALLOCN(aty); ALLOCN(xc);
    ALLOCM(y); ALLOCM(ys); ALLOCM(yr); ALLOCM(ya); ALLOCM(axb); ALLOCM(ax); ALLOCM(yc); ALLOCM(axc); ALLOCM(axp);
    if (Q) { ALLOCN(qx); ALLOCN(qxc); ALLOCN(qxa); ALLOCN(cq); }
#undef ALLOCN
#undef ALLOCM
    cost_norm = finite_norm(c, n);
    bound_norm = sqrt(finite_norm(rl, m) * finite_norm(rl, m) + finite_norm(ru, m) * finite_norm(ru, m));
    norm = orc_power_norm(A, P->power_iterations);
    st->matrix_norm = norm;
    eta_global = 0.9 / norm;
    eta_min = P->eta_min_factor / norm;
    eta_max = P->eta_max_factor / norm;
    if (P->primal_weight > 0.0)   /* an addition: an explicit starting weight */
        omega = fmin(1e6, fmax(1e-6, P->primal_weight));
    else if (cost_norm > 0.0 && bound_norm > 0.0)
        omega = fmin(1e6, fmax(1e-6, cost_norm / bound_norm));
    /* warm start (an addition: the reference always starts at zero) */
    if (x0) for (j = 0; j < n; ++j) x[j] = xb[j] = xr[j] = x0[j];
    if (y0) for (i = 0; i < m; ++i) y[i] = yr[i] = y0[i];
    if (x0) { orc_csr_mv(A, x, ax); memcpy(axp, ax, (size_t)m * sizeof(double)); }
    if (x0 && Q) orc_csr_mv(Q, x, qx);
    eta = eta_global;

    while (iteration < P->max_iterations) {
        const int period = P->restart_period < P->max_iterations - iteration
            ? P->restart_period : P->max_iterations - iteration;
        kkt_t cur, avg;
        double terms, scale, mu;
        int average_better, sufficient, necessary, artificial;
        if (period <= 0) break;
        for (k = 0; k < period; ++k) {
            if (P->adaptive_step) {
                const double sigma = eta * omega, tau = eta / omega;
                double dx2 = 0.0, dy2 = 0.0, inter = 0.0, dq = 0.0, movement, eta_bar, kk, shrink, grow;
                double eta_next;
                int accept;
                for (i = 0; i < m; ++i) {
                    const double axbar = 2.0 * ax[i] - axp[i];
                    const double v = y[i] + sigma * axbar;
                    yc[i] = v - sigma * clamp_to(v / sigma, rl[i], ru[i]);
                }
                orc_csr_mv(&At, yc, aty);
                /* QP: the gradient at x carries Qx */
                for (j = 0; j < n; ++j)
                    xc[j] = clamp_to(x[j] - tau * (c[j] + aty[j] + (Q ? qx[j] : 0.0)), l[j], u[j]);
                orc_csr_mv(A, xc, axc);
                if (Q) orc_csr_mv(Q, xc, qxc);
                for (j = 0; j < n; ++j) {
                    const double d = xc[j] - x[j];
                    dx2 += d * d;
                    if (Q) dq += d * (qxc[j] - qx[j]);
                }
                for (i = 0; i < m; ++i) {
                    const double dy = yc[i] - y[i], adx = axc[i] - ax[i];
                    dy2 += dy * dy; inter += dy * adx;
                }
                /* the local bound also covers dx'Q dx/2 (>= 0 for convex Q) */
                inter = fabs(inter) + 0.5 * fmax(dq, 0.0);
                movement = 0.5 * (omega * dx2 + dy2 / omega);
                eta_bar = (inter > 0.0 && movement > 0.0) ? movement / inter : eta_max;
                accept = eta <= eta_bar;
                kk = updates + 1.0;
                shrink = 1.0 - pow(kk + 1.0, -0.3);
                grow = 1.0 + pow(kk + 1.0, -0.6);
                eta_next = fmin(shrink * eta_bar, grow * eta);
                if (!(eta_next > 0.0) || !isfinite(eta_next)) eta_next = eta;
                eta_next = fmin(fmax(eta_next, eta_min), eta_max);
                eta = eta_next;
                updates = kk;
                if (accept) {
                    accept_count += 1.0;
                    for (j = 0; j < n; ++j) { x[j] = xc[j]; xs[j] += xc[j]; }
                    for (i = 0; i < m; ++i) { y[i] = yc[i]; ys[i] += yc[i]; axp[i] = ax[i]; ax[i] = axc[i]; }
                    if (Q) memcpy(qx, qxc, (size_t)n * sizeof(double));
                }
            } else {
                const double tau = eta_global / omega, sigma = eta_global * omega;
                orc_csr_mv(A, xb, axb);
                for (i = 0; i < m; ++i) {
                    const double v = y[i] + sigma * axb[i];
                    const double yi = v - sigma * clamp_to(v / sigma, rl[i], ru[i]);
                    y[i] = yi; ys[i] += yi;
                }
                orc_csr_mv(&At, y, aty);
                for (j = 0; j < n; ++j) {
                    const double xj = x[j];
                    const double xn = clamp_to(xj - tau * (c[j] + aty[j]), l[j], u[j]);
                    xb[j] = 2.0 * xn - xj;
                    x[j] = xn;
                    xs[j] += xn;
                }
            }
        }
        iteration += period;
        its_since_restart += period;
        ++windows_since_restart;
        st->host_syncs++;

        orc_csr_mv(A, x, ax);
        orc_csr_mv(&At, y, aty);
        if (Q) orc_csr_mv(Q, x, qx);
        cur = residuals_q(n, m, x, y, ax, aty, qx, c, l, u, rl, ru, xr, yr, cost_norm, bound_norm, cq);
        if (P->adaptive_step) {
            terms = accept_count - accepted_at_restart;
            accepted_total = accept_count;
            if (terms < 1.0) terms = 1.0;
        } else {
            terms = (double)windows_since_restart * (double)P->restart_period;
        }
        scale = 1.0 / terms;
        for (j = 0; j < n; ++j) xa[j] = scale * xs[j];
        for (i = 0; i < m; ++i) ya[i] = scale * ys[i];
        {
            double *scratch = P->adaptive_step ? axb : ax;
            orc_csr_mv(A, xa, scratch);
            orc_csr_mv(&At, ya, aty);
            if (Q) orc_csr_mv(Q, xa, qxa);
            avg = residuals_q(n, m, xa, ya, scratch, aty, qxa, c, l, u, rl, ru, xr, yr, cost_norm, bound_norm, cq);
        }
        st->host_syncs++;
        average_better = kkt_worst(&avg) < kkt_worst(&cur);
        best = average_better ? avg : cur;
        st->restart_merit[0] = kkt_worst(&cur);
        st->restart_merit[1] = kkt_worst(&avg);
        if (P->restart_rule == SK_GPU_RESTART_NDG) {
            /* the restart merits (only): NDG of the current and the
               averaged point against the last restart point */
            double *g2 = (double *)malloc(sizeof(double) * ((size_t)n + 1));
            double *a2 = (double *)malloc(sizeof(double) * ((size_t)m + 1));
            orc_csr_mv(A, x, a2);
            orc_csr_mv(&At, y, g2);
            for (j = 0; j < n; ++j) g2[j] += c[j];
            st->restart_merit[0] = orc_ndg(n, m, x, y, a2, g2, l, u, rl, ru, xr, yr, omega);
            orc_csr_mv(A, xa, a2);
            orc_csr_mv(&At, ya, g2);
            for (j = 0; j < n; ++j) g2[j] += c[j];
            st->restart_merit[1] = orc_ndg(n, m, xa, ya, a2, g2, l, u, rl, ru, xr, yr, omega);
            free(g2); free(a2);
        }
        if (best.primal <= P->eps_optimal && best.dual <= P->eps_optimal && best.gap <= P->eps_optimal) {
            if (average_better) {
                memcpy(x, xa, (size_t)n * sizeof(double));
                memcpy(y, ya, (size_t)m * sizeof(double));
            }
            st->converged = 1;
            break;
        }
        if (P->time_limit_seconds > 0.0 && wall_seconds() - t0 > P->time_limit_seconds) break;
        mu = kkt_worst(&best);
        if (P->restart_rule == SK_GPU_RESTART_NDG) {
            /* restart to the point of lower NDG, judged by its NDG */
            average_better = st->restart_merit[1] < st->restart_merit[0];
            best = average_better ? avg : cur;
            mu = st->restart_merit[average_better ? 1 : 0];
        }
        sufficient = mu <= 0.2 * mu_at_restart;
        necessary = mu <= 0.8 * mu_at_restart && mu > mu_previous;
        artificial = (double)its_since_restart >= 0.36 * (double)iteration;
        mu_previous = mu;
        if (P->verbosity > 1)   /* window trace for diagnosis */
            fprintf(stderr, "win %d cur %.4e avg %.4e p %.4e g %.4e pobj %.10e dobj %.10e omega %.4e eta %.4e "
                    "pmove %.4e dmove %.4e restart %d%d%d avg %d\n", iteration, kkt_worst(&cur), kkt_worst(&avg),
                    best.primal, best.gap, best.pobj, best.dobj, omega, eta, best.pmove, best.dmove, sufficient,
                    necessary, artificial, average_better);
        if (!(sufficient || necessary || artificial)) continue;
        if (average_better) {
            memcpy(x, xa, (size_t)n * sizeof(double));
            memcpy(y, ya, (size_t)m * sizeof(double));
            memcpy(xb, xa, (size_t)n * sizeof(double));
            if (P->adaptive_step) memcpy(ax, axb, (size_t)m * sizeof(double));
            if (Q) memcpy(qx, qxa, (size_t)n * sizeof(double));
        } else {
            memcpy(xb, x, (size_t)n * sizeof(double));
        }
        if (P->adaptive_step) {
            memcpy(axp, ax, (size_t)m * sizeof(double));
            accepted_at_restart = accepted_total;
        }
        memset(xs, 0, (size_t)n * sizeof(double));
        memset(ys, 0, (size_t)m * sizeof(double));
        windows_since_restart = 0;
        its_since_restart = 0;
        mu_at_restart = mu;
        mu_previous = INFINITY;
        st->restarts++;
        if (best.pmove > 1e-14 && best.dmove > 1e-14) {
            const double target = best.dmove / best.pmove;
            const double updated = exp(0.5 * log(target) + 0.5 * log(omega));
            if (isfinite(updated)) omega = fmin(1e6, fmax(1e-6, updated));
        }
        memcpy(xr, x, (size_t)n * sizeof(double));
        memcpy(yr, y, (size_t)m * sizeof(double));
    }
    memcpy(x_out, x, (size_t)n * sizeof(double));
    memcpy(y_out, y, (size_t)m * sizeof(double));
    st->iterations = iteration;
    st->fp64_iterations = iteration;
    st->primal_objective = best.pobj;
    st->dual_objective = best.dobj;
    st->relative_primal_residual = best.primal;
    st->relative_dual_residual = best.dual;
    st->relative_gap = best.gap;
    st->primal_weight = omega;
    if (P->adaptive_step) {
        st->final_step_size = eta;
        st->step_size_ratio = eta * norm;
        st->rejected_steps = (int)(updates - accept_count > 0.0 ? updates - accept_count : 0.0);
    } else {
        st->final_step_size = eta_global;
        st->step_size_ratio = 0.9;
    }
    st->seconds = wall_seconds() - t0;
    free(x); free(xb); free(xs); free(xr); free(xa); free(aty); free(xc);
    free(y); free(ys); free(yr); free(ya); free(axb); free(ax); free(yc); free(axc); free(axp);
    free(qx); free(qxc); free(qxa); free(cq);
    orc_csr_free(&At);
    return 0;
}

int orc_pdlp_solve(const orc_csr *A, const double *c, const double *l, const double *u,
                   const double *rl, const double *ru, const sk_gpu_pdlp_params *P,
                   const double *x0, const double *y0, double *x_out, double *y_out,
                   sk_gpu_pdlp_stats *st)
{
    return pdlp_solve(A, NULL, c, l, u, rl, ru, P, x0, y0, x_out, y_out, st);
}

int orc_pdlp_solve_qp(const orc_csr *A, const orc_csr *Q, const double *c, const double *l,
                      const double *u, const double *rl, const double *ru,
                      const sk_gpu_pdlp_params *P, const double *x0, const double *y0,
                      double *x_out, double *y_out, sk_gpu_pdlp_stats *st)
{
    return pdlp_solve(A, Q, c, l, u, rl, ru, P, x0, y0, x_out, y_out, st);
}

static double to_ieee(double b)
{
    if (b >= SK_INFINITY) return INFINITY;
    if (b <= -SK_INFINITY) return -INFINITY;
    return b;
}

/* run_first_order, src/lp/LpSolver.cpp:62-99: Ruiz, then Pock-Chambolle on
 * the Ruiz-scaled matrix, composed and applied once; x = C x', y = R y'. */
static int first_order(const orc_csr *A, const orc_csr *Q, const double *c, const double *l,
                       const double *u, const double *rl, const double *ru,
                       const sk_gpu_pdlp_params *P, double *x, double *y, sk_gpu_pdlp_stats *st)
{
    const int m = A->rows, n = A->cols;
    double *R1 = (double *)malloc((size_t)(m + 1) * sizeof(double));
    double *C1 = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *R2 = (double *)malloc((size_t)(m + 1) * sizeof(double));
    double *C2 = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *cs = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *ls = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *us = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *rls = (double *)malloc((size_t)(m + 1) * sizeof(double));
    double *rus = (double *)malloc((size_t)(m + 1) * sizeof(double));
    orc_csr S;
    int i, j, k, rc;
    S.rows = m; S.cols = n; S.nnz = A->nnz; S.p = A->p; S.j = A->j;
    S.v = (double *)malloc((size_t)(A->nnz + 1) * sizeof(double));
    if (P->scaling) {
        orc_ruiz(A, P->ruiz_iterations, P->ruiz_tol, R1, C1);
        for (i = 0; i < m; ++i)
            for (k = A->p[i]; k < A->p[i + 1]; ++k) S.v[k] = A->v[k] * R1[i] * C1[A->j[k]];
        orc_pock_chambolle(&S, 1.0, R2, C2);
        for (i = 0; i < m; ++i) R1[i] *= R2[i];
        for (j = 0; j < n; ++j) C1[j] *= C2[j];
    } else {
        for (i = 0; i < m; ++i) R1[i] = 1.0;
        for (j = 0; j < n; ++j) C1[j] = 1.0;
    }
    for (i = 0; i < m; ++i)
        for (k = A->p[i]; k < A->p[i + 1]; ++k) S.v[k] = A->v[k] * R1[i] * C1[A->j[k]];
    for (j = 0; j < n; ++j) {
        cs[j] = c[j] * C1[j];
        ls[j] = to_ieee(l[j]) / C1[j];
        us[j] = to_ieee(u[j]) / C1[j];
    }
    for (i = 0; i < m; ++i) { rls[i] = R1[i] * to_ieee(rl[i]); rus[i] = R1[i] * to_ieee(ru[i]); }
    if (Q) {
        /* the scaled Hessian C Q C (the row/column scaling comes from A alone) */
        orc_csr Qs;
        Qs.rows = Q->rows; Qs.cols = Q->cols; Qs.nnz = Q->nnz; Qs.p = Q->p; Qs.j = Q->j;
        Qs.v = (double *)malloc((size_t)(Q->nnz + 1) * sizeof(double));
        for (i = 0; i < n; ++i)
            for (k = Q->p[i]; k < Q->p[i + 1]; ++k) Qs.v[k] = Q->v[k] * C1[i] * C1[Q->j[k]];
        rc = pdlp_solve(&S, &Qs, cs, ls, us, rls, rus, P, NULL, NULL, x, y, st);
        free(Qs.v);
    } else {
        rc = pdlp_solve(&S, NULL, cs, ls, us, rls, rus, P, NULL, NULL, x, y, st);
    }
    for (j = 0; j < n; ++j) x[j] *= C1[j];
    for (i = 0; i < m; ++i) y[i] *= R1[i];
    free(R1); free(C1); free(R2); free(C2); free(cs); free(ls); free(us); free(rls); free(rus); free(S.v);
    return rc;
}

int orc_first_order(const orc_csr *A, const double *c, const double *l, const double *u,
                    const double *rl, const double *ru, const sk_gpu_pdlp_params *P,
                    double *x, double *y, sk_gpu_pdlp_stats *st)
{
    return first_order(A, NULL, c, l, u, rl, ru, P, x, y, st);
}

int orc_first_order_qp(const orc_csr *A, const orc_csr *Q, const double *c, const double *l,
                       const double *u, const double *rl, const double *ru,
                       const sk_gpu_pdlp_params *P, double *x, double *y, sk_gpu_pdlp_stats *st)
{
    if (!Q || Q->rows != A->cols || Q->cols != A->cols) return -1;
    return first_order(A, Q, c, l, u, rl, ru, P, x, y, st);
}

/* Relative KKT triple of (x, y) in the space the data is given in; bounds
 * with |b| >= 1e30 are absent.  Optional Q adds 1/2 x'Qx to the primal
 * objective, Qx to the reduced cost and -1/2 x'Qx to the dual objective. */
void orc_relative_kkt(const orc_csr *A, const orc_csr *Q, const double *c, const double *rl,
                      const double *ru, const double *l, const double *u, const double *x,
                      const double *y, double out[5])
{
    const int n = A->cols, m = A->rows;
    orc_csr At;
    double *ax = (double *)calloc((size_t)m + 1, sizeof(double));
    double *aty = (double *)calloc((size_t)n + 1, sizeof(double));
    double *qx = (double *)calloc((size_t)n + 1, sizeof(double));
    double *cc = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *ll = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *uu = (double *)malloc((size_t)(n + 1) * sizeof(double));
    double *rrl = (double *)malloc((size_t)(m + 1) * sizeof(double));
    double *rru = (double *)malloc((size_t)(m + 1) * sizeof(double));
    double *zx = (double *)calloc((size_t)n + 1, sizeof(double));
    double *zy = (double *)calloc((size_t)m + 1, sizeof(double));
    double xqx = 0.0, cost_norm, bound_norm;
    kkt_t k;
    int i, j;
    orc_csr_transpose(A, &At);
    orc_csr_mv(A, x, ax);
    orc_csr_mv(&At, y, aty);
    if (Q) orc_csr_mv(Q, x, qx);
    for (j = 0; j < n; ++j) {
        xqx += x[j] * qx[j];
        cc[j] = c[j] + qx[j];       /* reduced cost uses c + Qx */
        ll[j] = to_ieee(l[j]); uu[j] = to_ieee(u[j]);
    }
    for (i = 0; i < m; ++i) { rrl[i] = to_ieee(rl[i]); rru[i] = to_ieee(ru[i]); }
    cost_norm = finite_norm(c, n);
    bound_norm = sqrt(finite_norm(rrl, m) * finite_norm(rrl, m) + finite_norm(rru, m) * finite_norm(rru, m));
    k = residuals(n, m, x, y, ax, aty, cc, ll, uu, rrl, rru, zx, zy, cost_norm, bound_norm);
    /* residuals() used c+Qx as the cost: correct both objectives for Q */
    k.pobj = k.pobj - 0.5 * xqx;   /* (c+Qx)'x - 1/2 x'Qx = c'x + 1/2 x'Qx */
    k.dobj = k.dobj - 0.5 * xqx;
    k.gap = fabs(k.pobj - k.dobj) / (1.0 + fabs(k.pobj) + fabs(k.dobj));
    if (!isfinite(k.gap)) k.gap = INFINITY;
    out[0] = k.primal; out[1] = k.dual; out[2] = k.gap; out[3] = k.pobj; out[4] = k.dobj;
    free(ax); free(aty); free(qx); free(cc); free(ll); free(uu); free(rrl); free(rru); free(zx); free(zy);
    orc_csr_free(&At);
}
