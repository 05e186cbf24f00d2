/* LP presolve, postsolve and the first-order pipeline around the GPU engine.
 *
 * Presolve restates the reduction loop of src/lp/Presolve.cpp:69-572 on
 * the host: each pass runs, in this order, fixed columns, empty columns,
 * the row tests (crossed bounds, empty, activity infeasible, redundant,
 * forcing, singleton) and activity-based bound propagation, whose proposed
 * updates are collected per row and applied after the row.  Tolerances
 * are the reference's: 1e-7 for infeasibility, 1e-9 for reductions, a
 * 1e-7 relative improvement before a bound moves, a 1e-9 relative outward
 * relaxation of every derived bound, 1e-11 to call a column fixed.  The
 * reduced model is emitted in the reference's order and row encoding, so
 * it is the reference's bit for bit.  Presolve stays on the host: every
 * reduction reads the bounds the previous one left (a Gauss-Seidel sweep).
 * A row whose tests or propagation changed nothing, and none of whose
 * inputs changed since, is skipped: the result would again be no change,
 * so the outcome is unchanged while later passes touch only the rows near
 * the last reductions.
 *
 * sk_gpu_solve_lp is src/lp/LpSolver.cpp:127-333 with FIRST_ORDER:
 * presolve, the GPU PDLP solve (sk_gpu_pdlp_solve: Ruiz + Pock-Chambolle
 * scaling, restarted adaptive PDHG) of the reduced model, postsolve, the
 * objective from the original costs, and the original-space primal gate. */
#define _POSIX_C_SOURCE 200809L
#include "skg_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define INFEAS_TOL 1e-7
#define REDUCTION_TOL 1e-9
#define BOUND_IMPROVE_TOL 1e-7
#define BOUND_RELAX 1e-9
#define FIX_TOL 1e-11
#define FINAL_PRIMAL_TOL 1e-6      /* src/lp/LpSolver.cpp:15 */

typedef struct bounds { double lo, hi; } bounds;

/* Working state of one presolve call.  Rows are lo <= A x <= hi. */
typedef struct ps {
    int m, n;
    const int *rp, *rj;            /* CSR of A                              */
    const double *rv;
    hcsr csc;                      /* CSR of A^T: the columns of A          */
    const double *obj;
    double *row_lo, *row_hi;
    bounds *cb;                    /* column bounds, lo and hi side by side */
    char *row_active, *col_active;
    /* Change stamps.  Every change to a column (bounds, or fixed) takes the
     * next stamp; rt_seen[i] / pg_seen[i] is the first stamp after the start
     * of row i's last tests / propagation that changed nothing (0: none).
     * When no column of the row has a newer stamp, nothing the row reads
     * has changed (its bounds change only when one of its columns is
     * fixed), so redoing it would again change nothing and it is skipped. */
    unsigned *col_stamp, *rt_seen, *pg_seen, stamp;
    int *row_count, *col_count;
    double *fixed;
    char *removed;
    int infeasible, unbounded, pass;
    const char *reason;
    int reason_row, reason_col, reason_pass;
    /* per-row scratch: forcing pins and propagation proposals */
    int *buf_j, *low_j, *up_j;
    double *buf_v, *low_v, *up_v;
} ps;

typedef struct activity { double fmin, fmax; int nmin_inf, nmax_inf; } activity;

static double act_min(const activity *a) { return a->nmin_inf > 0 ? -INFINITY : a->fmin; }
static double act_max(const activity *a) { return a->nmax_inf > 0 ? INFINITY : a->fmax; }

/* Column j changed.  Should the stamps run out, every row is marked as
 * needing work again and counting restarts. */
static void touch_column(ps *P, int j)
{
    if (P->stamp == 0xfffffff0u) {
        memset(P->rt_seen, 0, sizeof(unsigned) * ((size_t)P->m + 1));
        memset(P->pg_seen, 0, sizeof(unsigned) * ((size_t)P->m + 1));
        memset(P->col_stamp, 0, sizeof(unsigned) * ((size_t)P->n + 1));
        P->stamp = 1;
    }
    P->col_stamp[j] = ++P->stamp;
}

/* Whether row i is unchanged since `seen` (see the stamps above). */
static int row_unchanged(const ps *P, int i, unsigned seen)
{
    int k;
    if (seen == 0) return 0;
    for (k = P->rp[i]; k < P->rp[i + 1]; ++k)
        if (P->rv[k] != 0.0 && P->col_stamp[P->rj[k]] >= seen) return 0;
    return 1;
}

/* Presolve.cpp:111-130: substitute the value into every active row. */
static void fix_column(ps *P, int j, double value)
{
    int k;
    if (!P->col_active[j]) return;
    P->col_active[j] = 0;
    P->removed[j] = 1;
    P->fixed[j] = value;
    touch_column(P, j);
    for (k = P->csc.p[j]; k < P->csc.p[j + 1]; ++k) {
        const int i = P->csc.j[k];
        const double a = P->csc.v[k];
        if (a == 0.0 || !P->row_active[i]) continue;
        if (isfinite(P->row_lo[i])) P->row_lo[i] -= a * value;
        if (isfinite(P->row_hi[i])) P->row_hi[i] -= a * value;
        --P->row_count[i];
    }
}

/* Presolve.cpp:132-144 */
static void drop_row(ps *P, int i)
{
    int k;
    if (!P->row_active[i]) return;
    P->row_active[i] = 0;
    for (k = P->rp[i]; k < P->rp[i + 1]; ++k) {
        const int j = P->rj[k];
        if (P->rv[k] == 0.0 || !P->col_active[j]) continue;
        --P->col_count[j];
    }
}

/* Presolve.cpp:155-177: activity bounds under the current column bounds,
 * computed afresh at every use (never cached across reductions). */
static activity row_activity(const ps *P, int i)
{
    activity act = {0.0, 0.0, 0, 0};
    int k;
    for (k = P->rp[i]; k < P->rp[i + 1]; ++k) {
        const int j = P->rj[k];
        const double a = P->rv[k];
        double mn, mx;
        if (a == 0.0 || !P->col_active[j]) continue;
        mn = a > 0.0 ? P->cb[j].lo : P->cb[j].hi;
        mx = a > 0.0 ? P->cb[j].hi : P->cb[j].lo;
        if (isfinite(mn)) act.fmin += a * mn; else ++act.nmin_inf;
        if (isfinite(mx)) act.fmax += a * mx; else ++act.nmax_inf;
    }
    return act;
}

/* Presolve.cpp:202-231: a derived bound is relaxed outward, taken only on a
 * relative improvement, and refused (never read as infeasibility) when it
 * would cross the opposite bound. */
static int tighten_upper(ps *P, int j, double limit)
{
    double relaxed, cur;
    if (!isfinite(limit)) return 0;
    relaxed = limit + BOUND_RELAX * (1.0 + fabs(limit));
    if (relaxed < P->cb[j].lo) return 0;
    cur = P->cb[j].hi;
    if (isfinite(cur) && !(relaxed < cur - BOUND_IMPROVE_TOL * (1.0 + fabs(cur)))) return 0;
    P->cb[j].hi = relaxed;
    touch_column(P, j);
    return 1;
}

static int tighten_lower(ps *P, int j, double limit)
{
    double relaxed, cur;
    if (!isfinite(limit)) return 0;
    relaxed = limit - BOUND_RELAX * (1.0 + fabs(limit));
    if (relaxed > P->cb[j].hi) return 0;
    cur = P->cb[j].lo;
    if (isfinite(cur) && !(relaxed > cur + BOUND_IMPROVE_TOL * (1.0 + fabs(cur)))) return 0;
    P->cb[j].lo = relaxed;
    touch_column(P, j);
    return 1;
}

static void fail(ps *P, const char *why, int row, int col)
{
    if (!P->reason) {
        P->reason = why;
        P->reason_row = row;
        P->reason_col = col;
        P->reason_pass = P->pass;
    }
    P->infeasible = 1;
}

/* Presolve.cpp:241-256 fixed columns; returns whether one was fixed. */
static int fixed_columns(ps *P)
{
    int j, changed = 0;
    for (j = 0; j < P->n; ++j) {
        const double lo = P->cb[j].lo, hi = P->cb[j].hi;
        if (!P->col_active[j]) continue;
        if (lo > hi + INFEAS_TOL * (1.0 + fabs(hi))) {
            fail(P, "column bounds crossed", -1, j);
            break;
        }
        if (isfinite(lo) && isfinite(hi) && hi - lo <= FIX_TOL) {
            fix_column(P, j, 0.5 * (lo + hi));
            changed = 1;
        }
    }
    return changed;
}

/* Presolve.cpp:259-286: a column in no active row goes to the bound its
 * cost prefers; an infinite preferred bound means unbounded. */
static int empty_columns(ps *P)
{
    int j, changed = 0;
    for (j = 0; j < P->n; ++j) {
        const double lo = P->cb[j].lo, hi = P->cb[j].hi, c = P->obj[j];
        double v;
        if (!P->col_active[j] || P->col_count[j] != 0) continue;
        if (c > 0.0) {
            if (!isfinite(lo)) { P->unbounded = 1; break; }
            v = lo;
        } else if (c < 0.0) {
            if (!isfinite(hi)) { P->unbounded = 1; break; }
            v = hi;
        } else {
            v = isfinite(lo) ? lo : (isfinite(hi) ? hi : 0.0);
        }
        fix_column(P, j, v);
        changed = 1;
    }
    return changed;
}

/* Presolve.cpp:366-390: pin every column of a forcing row at the bound that
 * gives the extreme activity (pins read before any is applied). */
static void force_row(ps *P, int i, int at_upper)
{
    int k, cnt = 0;
    for (k = P->rp[i]; k < P->rp[i + 1]; ++k) {
        const int j = P->rj[k];
        const double a = P->rv[k];
        double pin;
        if (a == 0.0 || !P->col_active[j]) continue;
        pin = (a > 0.0) == (at_upper != 0) ? P->cb[j].lo : P->cb[j].hi;
        if (!isfinite(pin)) continue;
        P->buf_j[cnt] = j;
        P->buf_v[cnt++] = pin;
    }
    for (k = 0; k < cnt; ++k) fix_column(P, P->buf_j[k], P->buf_v[k]);
    drop_row(P, i);
}

/* Presolve.cpp:394-434: a row with one active entry becomes bounds on that
 * column; the row is dropped only when both implied bounds were absorbed. */
static int singleton_row(ps *P, int i)
{
    int k, j = -1, took_lo, took_hi;
    double a = 0.0, ilo = -INFINITY, ihi = INFINITY, lo_b, hi_b;
    for (k = P->rp[i]; k < P->rp[i + 1]; ++k) {
        if (P->rv[k] == 0.0 || !P->col_active[P->rj[k]]) continue;
        j = P->rj[k];
        a = P->rv[k];
        break;
    }
    if (j < 0) return 0;
    lo_b = a > 0.0 ? P->row_lo[i] : P->row_hi[i];
    hi_b = a > 0.0 ? P->row_hi[i] : P->row_lo[i];
    if (isfinite(lo_b)) ilo = lo_b / a;
    if (isfinite(hi_b)) ihi = hi_b / a;
    took_lo = tighten_lower(P, j, ilo);
    took_hi = tighten_upper(P, j, ihi);
    if (!(took_lo || !isfinite(ilo) || ilo <= P->cb[j].lo + BOUND_RELAX * (1.0 + fabs(ilo)))) return 0;
    if (!(took_hi || !isfinite(ihi) || ihi >= P->cb[j].hi - BOUND_RELAX * (1.0 + fabs(ihi)))) return 0;
    drop_row(P, i);
    return 1;
}

/* Presolve.cpp:289-437: the row tests of one pass, in row order. */
static int row_tests(ps *P)
{
    int i, changed = 0;
    for (i = 0; i < P->m; ++i) {
        double lo, hi, scale, ascale, mn, mx;
        activity act;
        int at_upper, at_lower;
        if (!P->row_active[i] || row_unchanged(P, i, P->rt_seen[i])) continue;
        P->rt_seen[i] = P->stamp + 1;   /* stale again if this test changes a column */
        lo = P->row_lo[i];
        hi = P->row_hi[i];
        scale = 1.0 + fmax(isfinite(lo) ? fabs(lo) : 0.0, isfinite(hi) ? fabs(hi) : 0.0);
        if (lo > hi + INFEAS_TOL * scale) { fail(P, "row bounds crossed", i, -1); break; }
        if (P->row_count[i] == 0) {
            if (lo > INFEAS_TOL * scale || hi < -INFEAS_TOL * scale) {
                fail(P, "empty row violates its bounds", i, -1);
                break;
            }
            drop_row(P, i);
            changed = 1;
            continue;
        }
        act = row_activity(P, i);
        mn = act_min(&act);
        mx = act_max(&act);
        ascale = scale + fmax(act.nmin_inf ? 0.0 : fabs(act.fmin), act.nmax_inf ? 0.0 : fabs(act.fmax));
        if (mn > hi + INFEAS_TOL * ascale || mx < lo - INFEAS_TOL * ascale) {
            fail(P, "row activity cannot reach its bounds", i, -1);
            break;
        }
        if (mn >= lo - REDUCTION_TOL * scale && mx <= hi + REDUCTION_TOL * scale) {   /* redundant */
            drop_row(P, i);
            changed = 1;
            continue;
        }
        at_upper = act.nmin_inf == 0 && isfinite(hi) && fabs(act.fmin - hi) <= REDUCTION_TOL * scale;
        at_lower = act.nmax_inf == 0 && isfinite(lo) && fabs(act.fmax - lo) <= REDUCTION_TOL * scale;
        if (at_upper || at_lower) {
            force_row(P, i, at_upper);
            changed = 1;
            continue;
        }
        if (P->row_count[i] == 1) changed |= singleton_row(P, i);
    }
    return changed;
}

/* Presolve.cpp:466-524: for every column of every active row, the bound the
 * rest of the row's activity implies; proposals use the row-entry activity
 * and are applied after the row, lower bounds first. */
static int propagate(ps *P)
{
    int i, k, changed = 0;
    for (i = 0; i < P->m; ++i) {
        activity act;
        int nl = 0, nu = 0;
        const double lo = P->row_lo[i], hi = P->row_hi[i];
        if (!P->row_active[i] || row_unchanged(P, i, P->pg_seen[i])) continue;
        P->pg_seen[i] = P->stamp + 1;   /* stale again if this row moves a bound */
        act = row_activity(P, i);
        for (k = P->rp[i]; k < P->rp[i + 1]; ++k) {
            const int j = P->rj[k];
            const double a = P->rv[k];
            double mnb, mxb, rest_min = -INFINITY, rest_max = INFINITY;
            if (a == 0.0 || !P->col_active[j]) continue;
            mnb = a > 0.0 ? P->cb[j].lo : P->cb[j].hi;
            mxb = a > 0.0 ? P->cb[j].hi : P->cb[j].lo;
            if (act.nmin_inf == 0) rest_min = act.fmin - a * mnb;
            else if (act.nmin_inf == 1 && !isfinite(mnb)) rest_min = act.fmin;
            if (act.nmax_inf == 0) rest_max = act.fmax - a * mxb;
            else if (act.nmax_inf == 1 && !isfinite(mxb)) rest_max = act.fmax;
            if (isfinite(rest_min) && isfinite(hi)) {
                const double limit = (hi - rest_min) / a;
                if (a > 0.0) { P->up_j[nu] = j; P->up_v[nu++] = limit; }
                else { P->low_j[nl] = j; P->low_v[nl++] = limit; }
            }
            if (isfinite(rest_max) && isfinite(lo)) {
                const double limit = (lo - rest_max) / a;
                if (a > 0.0) { P->low_j[nl] = j; P->low_v[nl++] = limit; }
                else { P->up_j[nu] = j; P->up_v[nu++] = limit; }
            }
        }
        for (k = 0; k < nl; ++k) changed |= tighten_lower(P, P->low_j[k], P->low_v[k]);
        for (k = 0; k < nu; ++k) changed |= tighten_upper(P, P->up_j[k], P->up_v[k]);
    }
    return changed;
}

static int problem_check(const sk_gpu_lp_problem *p)
{
    int i, k;
    if (!p) return skg_fail("null problem");
    if (p->rows < 0 || p->cols < 0 || p->nnz < 0) return skg_fail("negative dimension");
    if (!p->row_ptr) return skg_fail("null row_ptr");
    if (p->nnz > 0 && (!p->col_idx || !p->values)) return skg_fail("null col_idx/values");
    if (p->cols > 0 && (!p->obj || !p->lower || !p->upper)) return skg_fail("null column vectors");
    if (p->rows > 0 && (!p->rhs || !p->slack_lower || !p->slack_upper)) return skg_fail("null row vectors");
    if (p->row_ptr[0] != 0 || p->row_ptr[p->rows] != p->nnz) return skg_fail("row_ptr must run from 0 to nnz");
    for (i = 0; i < p->rows; ++i)
        if (p->row_ptr[i + 1] < p->row_ptr[i]) return skg_fail("row_ptr must be non-decreasing");
    for (k = 0; k < p->nnz; ++k)
        if (p->col_idx[k] < 0 || p->col_idx[k] >= p->cols) return skg_fail("column index out of range");
    return 0;
}

void sk_gpu_lp_problem_free(sk_gpu_lp_problem *p)
{
    if (!p) return;
    free(p->row_ptr); free(p->col_idx); free(p->values); free(p->obj); free(p->rhs);
    free(p->row_types); free(p->lower); free(p->upper); free(p->slack_lower); free(p->slack_upper);
    memset(p, 0, sizeof(*p));
}

static int problem_alloc(sk_gpu_lp_problem *p, int rows, int cols, int nnz)
{
    memset(p, 0, sizeof(*p));
    p->rows = rows; p->cols = cols; p->nnz = nnz;
    p->row_ptr = (int *)calloc((size_t)rows + 1, sizeof(int));
    p->col_idx = (int *)malloc(sizeof(int) * ((size_t)nnz + 1));
    p->values = (double *)malloc(sizeof(double) * ((size_t)nnz + 1));
    p->obj = (double *)malloc(sizeof(double) * ((size_t)cols + 1));
    p->lower = (double *)malloc(sizeof(double) * ((size_t)cols + 1));
    p->upper = (double *)malloc(sizeof(double) * ((size_t)cols + 1));
    p->rhs = (double *)malloc(sizeof(double) * ((size_t)rows + 1));
    p->slack_lower = (double *)malloc(sizeof(double) * ((size_t)rows + 1));
    p->slack_upper = (double *)malloc(sizeof(double) * ((size_t)rows + 1));
    p->row_types = (char *)malloc((size_t)rows + 1);
    if (!p->row_ptr || !p->col_idx || !p->values || !p->obj || !p->lower || !p->upper || !p->rhs ||
        !p->slack_lower || !p->slack_upper || !p->row_types) {
        sk_gpu_lp_problem_free(p);
        return skg_fail("out of host memory");
    }
    return 0;
}

/* Presolve.cpp:535-569: lo <= A x <= hi re-encoded as A x + s = rhs, with
 * rhs at the finite upper bound, else the finite lower bound, else 0 (a
 * free row). */
static void encode_row(double lo, double hi, double *rhs, double *s_lo, double *s_hi, char *type)
{
    if (isfinite(hi)) {
        *rhs = hi;
        *s_lo = 0.0;
        *s_hi = isfinite(lo) ? hi - lo : INFINITY;
        *type = isfinite(lo) && hi - lo <= FIX_TOL ? 'E' : 'L';
    } else if (isfinite(lo)) {
        *rhs = lo;
        *s_lo = -INFINITY;
        *s_hi = 0.0;
        *type = 'G';
    } else {
        *rhs = 0.0;
        *s_lo = -INFINITY;
        *s_hi = INFINITY;
        *type = 'N';
    }
}

int sk_gpu_lp_problem_from_model(const sk_model *m, sk_gpu_lp_problem *p)
{
    hcsr A;
    int i, j;
    skg_clear_error();
    if (!m || !p) return skg_fail("null argument");
    if (sk_gpu_model_validate(m) != SK_OK) return skg_fail("invalid model");
    if (m->Q) return skg_fail("the LP pipeline takes linear programs only");
    if (hcsr_from_csc(&m->A, &A)) return skg_fail("out of host memory");
    if (problem_alloc(p, m->nrow, m->ncol, A.nnz)) { hcsr_free(&A); return -1; }
    memcpy(p->row_ptr, A.p, sizeof(int) * ((size_t)m->nrow + 1));
    if (A.nnz) {
        memcpy(p->col_idx, A.j, sizeof(int) * (size_t)A.nnz);
        memcpy(p->values, A.v, sizeof(double) * (size_t)A.nnz);
    }
    hcsr_free(&A);
    for (j = 0; j < m->ncol; ++j) {
        p->obj[j] = m->c[j];
        p->lower[j] = SK_IS_NEG_INF(m->clow[j]) ? -INFINITY : m->clow[j];
        p->upper[j] = SK_IS_INF(m->cupp[j]) ? INFINITY : m->cupp[j];
    }
    for (i = 0; i < m->nrow; ++i) {
        const double lo = SK_IS_NEG_INF(m->rlow[i]) ? -INFINITY : m->rlow[i];
        const double hi = SK_IS_INF(m->rupp[i]) ? INFINITY : m->rupp[i];
        encode_row(lo, hi, &p->rhs[i], &p->slack_lower[i], &p->slack_upper[i], &p->row_types[i]);
    }
    return 0;
}

void sk_gpu_presolve_free(sk_gpu_presolve_result *r)
{
    if (!r) return;
    sk_gpu_lp_problem_free(&r->reduced);
    free(r->kept_columns); free(r->kept_rows); free(r->fixed_value); free(r->column_removed);
    memset(r, 0, sizeof(*r));
}

static void ps_free(ps *P)
{
    hcsr_free(&P->csc);
    free(P->row_lo); free(P->row_hi); free(P->cb);
    free(P->row_active); free(P->col_active); free(P->row_count); free(P->col_count);
    free(P->col_stamp); free(P->rt_seen); free(P->pg_seen);
    free(P->buf_j); free(P->low_j); free(P->up_j); free(P->buf_v); free(P->low_v); free(P->up_v);
}

/* One reduced row as the reference's CSR assembly leaves it
 * (src/sparse/CSRMatrix.cpp:20-86): entries by ascending column,
 * duplicates summed in input order. */
typedef struct entry { int col, seq; double v; } entry;

static int entry_cmp(const void *a, const void *b)
{
    const entry *x = (const entry *)a, *y = (const entry *)b;
    if (x->col != y->col) return x->col < y->col ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

static int row_emit(const ps *P, int i, const int *newc, entry *buf, int *cj, double *cv)
{
    int k, q, cnt = 0, out = 0, sorted = 1;
    for (k = P->rp[i]; k < P->rp[i + 1]; ++k) {
        const int j = P->rj[k];
        if (P->rv[k] == 0.0 || !P->col_active[j]) continue;
        buf[cnt].col = newc[j];
        buf[cnt].seq = cnt;
        buf[cnt].v = P->rv[k];
        if (cnt > 0 && buf[cnt - 1].col >= buf[cnt].col) sorted = 0;
        ++cnt;
    }
    if (!sorted) qsort(buf, (size_t)cnt, sizeof(entry), entry_cmp);
    for (q = 0; q < cnt; ++q) {
        if (out > 0 && cj[out - 1] == buf[q].col) { cv[out - 1] += buf[q].v; continue; }
        cj[out] = buf[q].col;
        cv[out++] = buf[q].v;
    }
    return out;
}

/* src/lp/Presolve.cpp:69-572 presolve */
int sk_gpu_presolve(const sk_gpu_lp_problem *problem, int max_passes, sk_gpu_presolve_result *out)
{
    ps P;
    hcsr A;
    int i, j, k, pass, rn = 0, rm = 0, nnz = 0, maxrow = 1, *newc = NULL, rc = -1;
    entry *ebuf = NULL;
    skg_clear_error();
    if (!out) return skg_fail("null result");
    memset(out, 0, sizeof(*out));
    if (problem_check(problem)) return -1;
    if (max_passes <= 0) max_passes = 20;
    memset(&P, 0, sizeof(P));
    P.m = problem->rows;
    P.n = problem->cols;
    P.rp = problem->row_ptr; P.rj = problem->col_idx; P.rv = problem->values;
    P.obj = problem->obj;
    A.rows = P.m; A.cols = P.n; A.nnz = problem->nnz;
    A.p = problem->row_ptr; A.j = problem->col_idx; A.v = problem->values;
    out->original_n_rows = P.m;
    out->original_n_cols = P.n;
    out->reason_row = out->reason_col = out->reason_pass = -1;
    for (i = 0; i < P.m; ++i) if (P.rp[i + 1] - P.rp[i] > maxrow) maxrow = P.rp[i + 1] - P.rp[i];
    P.row_lo = (double *)malloc(sizeof(double) * ((size_t)P.m + 1));
    P.row_hi = (double *)malloc(sizeof(double) * ((size_t)P.m + 1));
    P.cb = (bounds *)malloc(sizeof(bounds) * ((size_t)P.n + 1));
    P.row_active = (char *)malloc((size_t)P.m + 1);
    P.col_active = (char *)malloc((size_t)P.n + 1);
    P.row_count = (int *)calloc((size_t)P.m + 1, sizeof(int));
    P.rt_seen = (unsigned *)calloc((size_t)P.m + 1, sizeof(unsigned));
    P.pg_seen = (unsigned *)calloc((size_t)P.m + 1, sizeof(unsigned));
    P.col_stamp = (unsigned *)calloc((size_t)P.n + 1, sizeof(unsigned));
    P.stamp = 1;
    P.col_count = (int *)calloc((size_t)P.n + 1, sizeof(int));
    P.buf_j = (int *)malloc(sizeof(int) * (size_t)maxrow);
    P.buf_v = (double *)malloc(sizeof(double) * (size_t)maxrow);
    P.low_j = (int *)malloc(sizeof(int) * 2 * (size_t)maxrow);
    P.low_v = (double *)malloc(sizeof(double) * 2 * (size_t)maxrow);
    P.up_j = (int *)malloc(sizeof(int) * 2 * (size_t)maxrow);
    P.up_v = (double *)malloc(sizeof(double) * 2 * (size_t)maxrow);
    out->fixed_value = (double *)calloc((size_t)P.n + 1, sizeof(double));
    out->column_removed = (char *)calloc((size_t)P.n + 1, 1);
    newc = (int *)malloc(sizeof(int) * ((size_t)P.n + 1));
    if (!P.row_lo || !P.row_hi || !P.cb || !P.row_active || !P.col_active ||
        !P.row_count || !P.col_count || !P.rt_seen || !P.pg_seen || !P.col_stamp || !P.buf_j || !P.buf_v || !P.low_j || !P.low_v || !P.up_j ||
        !P.up_v || !out->fixed_value || !out->column_removed || !newc || hcsr_transpose(&A, &P.csc)) {
        skg_fail("out of host memory");
        goto done;
    }
    P.fixed = out->fixed_value;
    P.removed = out->column_removed;
    for (i = 0; i < P.m; ++i) {
        P.row_lo[i] = problem->rhs[i] - problem->slack_upper[i];
        P.row_hi[i] = problem->rhs[i] - problem->slack_lower[i];
        P.row_active[i] = 1;
    }
    for (j = 0; j < P.n; ++j) {
        P.cb[j].lo = problem->lower[j];
        P.cb[j].hi = problem->upper[j];
        P.col_active[j] = 1;
    }
    for (i = 0; i < P.m; ++i)
        for (k = P.rp[i]; k < P.rp[i + 1]; ++k) {
            if (P.rv[k] == 0.0) continue;
            ++P.row_count[i];
            ++P.col_count[P.rj[k]];
        }

    for (pass = 0; pass < max_passes && !P.infeasible && !P.unbounded; ++pass) {
        int changed;
        P.pass = pass;
        changed = fixed_columns(&P);
        if (P.infeasible) break;
        changed |= empty_columns(&P);
        if (P.unbounded) break;
        changed |= row_tests(&P);
        if (P.infeasible) break;
        changed |= propagate(&P);
        if (!changed) break;
    }

    out->reason = P.reason;
    if (P.reason) {
        out->reason_row = P.reason_row;
        out->reason_col = P.reason_col;
        out->reason_pass = P.reason_pass;
    }
    if (P.infeasible || P.unbounded) {
        out->status = P.infeasible ? SK_GPU_PRESOLVE_INFEASIBLE : SK_GPU_PRESOLVE_UNBOUNDED;
        out->removed_rows = P.m;
        out->removed_cols = P.n;
        rc = 0;
        goto done;
    }

    /* the reduced model: kept columns and rows in original order */
    out->kept_columns = (int *)malloc(sizeof(int) * ((size_t)P.n + 1));
    out->kept_rows = (int *)malloc(sizeof(int) * ((size_t)P.m + 1));
    if (!out->kept_columns || !out->kept_rows) { skg_fail("out of host memory"); goto done; }
    for (j = 0; j < P.n; ++j) {
        newc[j] = -1;
        if (!P.col_active[j]) continue;
        newc[j] = rn;
        out->kept_columns[rn++] = j;
    }
    for (i = 0; i < P.m; ++i) {
        if (!P.row_active[i]) continue;
        out->kept_rows[rm++] = i;
        for (k = P.rp[i]; k < P.rp[i + 1]; ++k) if (P.rv[k] != 0.0 && P.col_active[P.rj[k]]) ++nnz;
    }
    out->removed_rows = P.m - rm;
    out->removed_cols = P.n - rn;
    if (problem_alloc(&out->reduced, rm, rn, nnz)) goto done;
    ebuf = (entry *)malloc(sizeof(entry) * (size_t)maxrow);
    if (!ebuf) { skg_fail("out of host memory"); goto done; }
    {
        sk_gpu_lp_problem *R = &out->reduced;
        int at = 0;
        for (k = 0; k < rm; ++k) {
            const int o = out->kept_rows[k];
            at += row_emit(&P, o, newc, ebuf, R->col_idx + at, R->values + at);
            R->row_ptr[k + 1] = at;
            encode_row(P.row_lo[o], P.row_hi[o], &R->rhs[k], &R->slack_lower[k], &R->slack_upper[k],
                       &R->row_types[k]);
        }
        R->nnz = at;
        for (k = 0; k < rn; ++k) {
            const int o = out->kept_columns[k];
            R->obj[k] = problem->obj[o];
            R->lower[k] = P.cb[o].lo;
            R->upper[k] = P.cb[o].hi;
        }
    }
    out->status = SK_GPU_PRESOLVE_OK;
    rc = 0;
done:
    ps_free(&P);
    free(newc);
    free(ebuf);
    if (rc) sk_gpu_presolve_free(out);
    return rc;
}

/* src/lp/Presolve.cpp:574-589 postsolve */
int sk_gpu_postsolve(const sk_gpu_presolve_result *r, const double *reduced_x, int reduced_len, double *x)
{
    int j, k, kept;
    if (!r || !x) return skg_fail("null argument");
    if (reduced_len > 0 && !reduced_x) return skg_fail("null reduced solution");
    kept = r->original_n_cols - r->removed_cols;
    for (j = 0; j < r->original_n_cols; ++j)
        x[j] = r->column_removed && r->column_removed[j] ? r->fixed_value[j] : 0.0;
    for (k = 0; k < kept && k < reduced_len; ++k) x[r->kept_columns[k]] = reduced_x[k];
    return 0;
}

/* src/lp/LpSolver.cpp:23-50 original_space_primal_residual */
double sk_gpu_lp_primal_residual(const sk_gpu_lp_problem *p, const double *x)
{
    double row_v = 0.0, bound_v = 0.0, rhs_norm = 0.0;
    int i, j, k;
    for (i = 0; i < p->rows; ++i) {
        double ax = 0.0;
        const double lo = p->rhs[i] - p->slack_upper[i], hi = p->rhs[i] - p->slack_lower[i];
        if (p->cols > 0)
            for (k = p->row_ptr[i]; k < p->row_ptr[i + 1]; ++k) ax += p->values[k] * x[p->col_idx[k]];
        if (isfinite(lo)) row_v = fmax(row_v, lo - ax);
        if (isfinite(hi)) row_v = fmax(row_v, ax - hi);
        rhs_norm = fmax(rhs_norm, fabs(p->rhs[i]));
    }
    for (j = 0; j < p->cols; ++j) {
        bound_v = fmax(bound_v, p->lower[j] - x[j]);
        bound_v = fmax(bound_v, x[j] - p->upper[j]);
    }
    row_v = fmax(0.0, row_v);
    bound_v = fmax(0.0, bound_v);
    return fmax(row_v / (1.0 + rhs_norm), bound_v);
}

void sk_gpu_lp_solver_options_default(sk_gpu_lp_solver_options *o)
{
    memset(o, 0, sizeof(*o));
    o->use_presolve = 1;
    o->presolve_passes = 20;
    sk_gpu_pdlp_params_default(&o->pdlp);
}

/* src/lp/LpSolver.cpp:62-99 run_first_order on the GPU engine: the rows go
 * in as rhs - slack_upper <= A x <= rhs - slack_lower. */
static int first_order(const sk_gpu_lp_problem *t, const sk_gpu_pdlp_params *params, double *x,
                       sk_gpu_pdlp_stats *stats)
{
    const int m = t->rows, n = t->cols;
    double *rl = (double *)malloc(sizeof(double) * ((size_t)m + 1));
    double *ru = (double *)malloc(sizeof(double) * ((size_t)m + 1));
    double *y = (double *)malloc(sizeof(double) * ((size_t)m + 1));
    int i, rc;
    if (!rl || !ru || !y) { free(rl); free(ru); free(y); return skg_fail("out of host memory"); }
    for (i = 0; i < m; ++i) {
        rl[i] = t->rhs[i] - t->slack_upper[i];
        ru[i] = t->rhs[i] - t->slack_lower[i];
    }
    rc = sk_gpu_pdlp_solve(m, n, t->nnz, t->row_ptr, t->col_idx, t->values, NULL, NULL, NULL, 0, t->obj, rl,
                           ru, t->lower, t->upper, NULL, NULL, params, x, y, stats);
    free(rl); free(ru); free(y);
    return rc;
}

/* src/lp/LpSolver.cpp:127-333 solve_lp, method FIRST_ORDER */
int sk_gpu_solve_lp(const sk_gpu_lp_problem *problem, const sk_gpu_lp_solver_options *options, double *x,
                    sk_gpu_lp_solution *sol)
{
    sk_gpu_lp_solver_options od;
    sk_gpu_presolve_result red;
    const sk_gpu_lp_problem *target = problem;
    double *rx = NULL, t0, t1;
    int j, rc;
    skg_clear_error();
    if (!sol || (problem && problem->cols > 0 && !x)) return skg_fail("null argument");
    memset(sol, 0, sizeof(*sol));
    memset(&red, 0, sizeof(red));
    sol->status = SK_GPU_LP_NUMERICAL_FAILURE;
    if (problem_check(problem)) return -1;
    if (!options) { sk_gpu_lp_solver_options_default(&od); options = &od; }
    if (options->use_presolve) {
        t0 = skg_now();
        if (sk_gpu_presolve(problem, options->presolve_passes, &red)) return -1;
        sol->presolve_seconds = skg_now() - t0;
        if (red.status != SK_GPU_PRESOLVE_OK) {
            sol->status = red.status == SK_GPU_PRESOLVE_INFEASIBLE ? SK_GPU_LP_INFEASIBLE : SK_GPU_LP_UNBOUNDED;
            sk_gpu_presolve_free(&red);
            return 0;
        }
        sol->presolve_removed_rows = red.removed_rows;
        sol->presolve_removed_cols = red.removed_cols;
        target = &red.reduced;
    }
    t1 = skg_now();
    rx = (double *)calloc((size_t)target->cols + 1, sizeof(double));
    if (!rx) { sk_gpu_presolve_free(&red); return skg_fail("out of host memory"); }
    if (target->cols == 0) {
        sol->status = SK_GPU_LP_OPTIMAL;   /* presolve fixed every column */
    } else {
        sol->used_first_order = 1;
        rc = first_order(target, &options->pdlp, rx, &sol->pdlp);
        if (rc == -1) { free(rx); sk_gpu_presolve_free(&red); return -1; }
        sol->iterations = sol->pdlp.iterations;
        /* a non-finite iterate (-2) plays the reference's caught exception */
        sol->status = rc == 0 && sol->pdlp.converged ? SK_GPU_LP_OPTIMAL : SK_GPU_LP_ITERATION_LIMIT;
    }
    sol->solve_seconds = skg_now() - t1;
    if (sol->status == SK_GPU_LP_OPTIMAL) {
        if (options->use_presolve) sk_gpu_postsolve(&red, rx, target->cols, x);
        else if (problem->cols > 0) memcpy(x, rx, sizeof(double) * (size_t)problem->cols);
        sol->has_x = 1;
        sol->objective_value = 0.0;
        for (j = 0; j < problem->cols; ++j) sol->objective_value += problem->obj[j] * x[j];
        sol->primal_residual = sk_gpu_lp_primal_residual(problem, x);
        if (!(sol->primal_residual <= FINAL_PRIMAL_TOL)) sol->status = SK_GPU_LP_NUMERICAL_FAILURE;
    }
    free(rx);
    sk_gpu_presolve_free(&red);
    return 0;
}
