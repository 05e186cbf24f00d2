/* Restarted-average PDHG (PDLP) control loop on the host, C11.
 *
 * Algorithm spec: src/cuda/GpuPdlp.cpp:252-534 (windows of restart_period
 * iterations with no host synchronisation inside, evaluation of the current
 * and the averaged iterate, restart criteria 0.2 / 0.8 / 0.36, primal-weight
 * log-smoothing) and src/lp/LpSolver.cpp:62-99 (Ruiz then Pock-Chambolle
 * scaling, unscaling x = C x').  Additions for MI355X: the problem is
 * permuted so that every row-length bin of A and of A^T is a contiguous row
 * range; FP32 iterates with FP64 checks and an FP64 tail; optional hipGraph
 * capture of a whole restart window; warm start. */
#define _POSIX_C_SOURCE 200809L
#include "skg_internal.h"

#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

skg_tuning g_skg_tuning = {4, 512};

/* Block partials (A, A^T and Q products together) up to which the step
 * decision is fused into the next dual prox. */
#define FUSE_MAX_PARTIALS SKK_FUSE_PARTIALS

void sk_gpu_pdlp_params_default(sk_gpu_pdlp_params *p)
{
    memset(p, 0, sizeof(*p));
    p->eps_optimal = 1e-6;
    p->max_iterations = 500000;
    p->restart_period = 64;
    p->time_limit_seconds = 0.0;
    p->power_iterations = 30;
    p->adaptive_step = 1;
    p->eta_min_factor = 1e-6;
    p->eta_max_factor = 1e4;
    p->precision = SK_GPU_PREC_FP64;
    p->tail_switch = 1e-5;
    p->tail_stall_windows = 32;
    p->scaling = 1;
    p->ruiz_iterations = 100;
    p->ruiz_tol = 1e-2;
    p->use_graph = 0;
    p->verbosity = 0;
    p->refine = 1;
    p->cancel = NULL;
    p->restart_rule = SK_GPU_RESTART_KKT;
}

static double to_ieee(double b)
{
    if (b >= SK_INFINITY) return INFINITY;
    if (b <= -SK_INFINITY) return -INFINITY;
    return b;
}

static double finite_norm(const double *v, int n)       /* GpuPdlp.cpp:84-90 */
{
    double acc = 0.0;
    int k;
    for (k = 0; k < n; ++k) if (isfinite(v[k])) acc += v[k] * v[k];
    return sqrt(acc);
}

/* ---------------------------------------------------------------- scaling */

/* Ruiz equilibration, src/lp/Scaling.cpp:11-92: simultaneous row/column
 * division by sqrt of the current inf-norms until every nonzero norm is
 * within tol of 1; identity if that does not happen within max_iters.
 * Rows and columns are swept by host threads; every value is a maximum or
 * an elementwise update, so the result does not depend on the thread
 * count. */
void skg_ruiz(const hcsr *a, const hcsr *at, int max_iters, double tol, double *R, double *C)
{
    const int m = a->rows, n = a->cols;
    int i, j, iter, converged = 0;
    double *rn, *cn;
    for (i = 0; i < m; ++i) R[i] = 1.0;
    for (j = 0; j < n; ++j) C[j] = 1.0;
    if (a->nnz == 0 || m == 0 || n == 0) return;
    rn = (double *)malloc(sizeof(double) * (size_t)m);
    cn = (double *)malloc(sizeof(double) * (size_t)n);
    if (!rn || !cn) { free(rn); free(cn); return; }
    for (iter = 0; iter < max_iters; ++iter) {
        double dev = 0.0;
#pragma omp parallel for schedule(dynamic, 8192) reduction(max : dev)
        for (i = 0; i < m; ++i) {
            double best = 0.0;
            int k;
            for (k = a->p[i]; k < a->p[i + 1]; ++k) {
                const double v = fabs(a->v[k]) * C[a->j[k]];
                if (v > best) best = v;
            }
            rn[i] = R[i] * best;
            if (rn[i] > 0.0 && fabs(rn[i] - 1.0) > dev) dev = fabs(rn[i] - 1.0);
        }
#pragma omp parallel for schedule(dynamic, 8192) reduction(max : dev)
        for (j = 0; j < n; ++j) {
            double best = 0.0;
            int k;
            for (k = at->p[j]; k < at->p[j + 1]; ++k) {
                const double v = fabs(at->v[k]) * R[at->j[k]];
                if (v > best) best = v;
            }
            cn[j] = C[j] * best;
            if (cn[j] > 0.0 && fabs(cn[j] - 1.0) > dev) dev = fabs(cn[j] - 1.0);
        }
        if (dev < tol) { converged = 1; break; }
#pragma omp parallel for schedule(static)
        for (i = 0; i < m; ++i) if (rn[i] > 0.0) R[i] /= sqrt(rn[i]);
#pragma omp parallel for schedule(static)
        for (j = 0; j < n; ++j) if (cn[j] > 0.0) C[j] /= sqrt(cn[j]);
    }
    if (!converged) {
        for (i = 0; i < m; ++i) R[i] = 1.0;
        for (j = 0; j < n; ++j) C[j] = 1.0;
    }
    free(rn); free(cn);
}

/* Pock-Chambolle with alpha = 1 on the matrix a scaled by (R0, C0),
 * src/lp/Scaling.cpp:111-145: R_i = 1/sqrt(sum_j |a'_ij|),
 * C_j = 1/sqrt(sum_i |a'_ij|); empty rows/columns keep 1.  Multiplies the
 * result into R0, C0 (compose_scaling, Scaling.cpp:147-152).  Column sums
 * run over at = a^T, whose rows list each column's entries in increasing
 * row order -- the order of a row-major sweep -- so threads change no bit. */
void skg_pock_chambolle_into(const hcsr *a, const hcsr *at, double *R, double *C)
{
    const int m = a->rows, n = a->cols;
    int i, j;
    double *ra = (double *)calloc((size_t)m + 1, sizeof(double));
    double *ca = (double *)calloc((size_t)n + 1, sizeof(double));
    if (!ra || !ca) { free(ra); free(ca); return; }
#pragma omp parallel for schedule(dynamic, 8192)
    for (i = 0; i < m; ++i) {
        int k;
        for (k = a->p[i]; k < a->p[i + 1]; ++k) {
            const double v = fabs(a->v[k] * R[i] * C[a->j[k]]);
            if (v != 0.0) ra[i] += v;
        }
    }
#pragma omp parallel for schedule(dynamic, 8192)
    for (j = 0; j < n; ++j) {
        int k;
        for (k = at->p[j]; k < at->p[j + 1]; ++k) {
            const int r = at->j[k];
            const double v = fabs(at->v[k] * R[r] * C[j]);
            if (v != 0.0) ca[j] += v;
        }
    }
    for (i = 0; i < m; ++i) if (ra[i] > 0.0) R[i] *= 1.0 / sqrt(ra[i]);
    for (j = 0; j < n; ++j) if (ca[j] > 0.0) C[j] *= 1.0 / sqrt(ca[j]);
    free(ra); free(ca);
}

/* ----------------------------------------------------------------- engine */

typedef struct engine {
    int m, n, mixed;
    dcsr A, AT, Q;
    int has_q;
    int *perm_r, *perm_c;        /* slot -> original index */
    double *R, *C;               /* scale factors, original indexing */
    double cost_norm, bound_norm;
    void *dev[96];               /* every device allocation, freed together */
    int ndev;
    skk_pdhg P;
    double *pv, *pw, *pv2, *pnorm;
    double *h_eval;              /* pinned: 16 doubles + state */
    skk_state *h_state;
    void *graph[2];
    int graph_period[2];
    skk_unpermute U;             /* slot -> original order on the device */
    long long launches;
    int *rp_host[3];             /* host row offsets of A, A^T, Q (autotune) */
    int tile_nnz, grid_cap, probe_cap, xcache, stream, tuned;   /* tiling of the products */
    double tune_seconds;
    int unfused;
} engine;

static int dalloc(engine *E, void **p, size_t bytes)
{
    int rc = skk_malloc(p, bytes);
    if (rc) return skg_hip_fail("hipMalloc", rc);
    if (E->ndev < 96) E->dev[E->ndev++] = *p;
    else { skk_free(*p); *p = NULL; return skg_fail("too many device allocations"); }
    return 0;
}

static void engine_free(engine *E)
{
    int k;
    for (k = 0; k < 2; ++k) skk_graph_destroy(E->graph[k]);
    for (k = 0; k < E->ndev; ++k) skk_free(E->dev[k]);
    dcsr_free(&E->A); dcsr_free(&E->AT); dcsr_free(&E->Q);
    skk_host_free(E->h_eval);
    free(E->perm_r); free(E->perm_c); free(E->R); free(E->C);
    for (k = 0; k < 3; ++k) free(E->rp_host[k]);
    memset(E, 0, sizeof(*E));
}

/* sum64: the running sums of accepted iterates are FP64 whatever es is. */
static int alloc_iter(engine *E, skk_iter *it, size_t es, int sum64)
{
    const size_t n = (size_t)(E->n ? E->n : 1), m = (size_t)(E->m ? E->m : 1);
    int k, rc;
    it->sum64 = sum64;
    for (k = 0; k < 2; ++k) {
        if ((rc = dalloc(E, &it->X[k], es * n))) return rc;
        if ((rc = dalloc(E, &it->Y[k], es * m))) return rc;
    }
    for (k = 0; k < 3; ++k) if ((rc = dalloc(E, &it->AX[k], es * m))) return rc;
    const size_t ss = sum64 ? sizeof(double) : es;   /* running-sum element */
    if ((rc = dalloc(E, &it->xsum, ss * n)) || (rc = dalloc(E, &it->ysum, ss * m)) ||
        (rc = dalloc(E, &it->xr, es * n)) || (rc = dalloc(E, &it->yr, es * m)) ||
        (rc = dalloc(E, &it->xavg, es * n)) || (rc = dalloc(E, &it->yavg, es * m)) ||
        (rc = dalloc(E, &it->xbar, es * n)))
        return rc;
    if (E->has_q && ((rc = dalloc(E, &it->QX[0], es * n)) || (rc = dalloc(E, &it->QX[1], es * n))))
        return rc;
    return 0;
}

/* Scale (original -> slot order) and upload c, bounds in FP64 (and FP32). */
static int engine_set_data(engine *E, const double *c, const double *rl, const double *ru,
                           const double *l, const double *u)
{
    const int m = E->m, n = E->n;
    double *tc = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *tl = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *tu = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    double *trl = (double *)malloc(sizeof(double) * (size_t)(m + 1));
    double *tru = (double *)malloc(sizeof(double) * (size_t)(m + 1));
    int k, rc = 0;
    skk_pdhg *Q = &E->P;
    if (!tc || !tl || !tu || !trl || !tru) { rc = skg_fail("out of host memory"); goto done; }
    for (k = 0; k < n; ++k) {
        const int col = E->perm_c[k];
        tc[k] = c[col] * E->C[col];
        tl[k] = to_ieee(l[col]) / E->C[col];
        tu[k] = to_ieee(u[col]) / E->C[col];
    }
    for (k = 0; k < m; ++k) {
        const int r = E->perm_r[k];
        trl[k] = E->R[r] * to_ieee(rl[r]);
        tru[k] = E->R[r] * to_ieee(ru[r]);
    }
    E->cost_norm = finite_norm(tc, n);
    E->bound_norm = sqrt(finite_norm(trl, m) * finite_norm(trl, m) + finite_norm(tru, m) * finite_norm(tru, m));
#define PUT_D(dst, src, cnt) if (!rc && (rc = skk_h2d((void *)(dst), (src), sizeof(double) * (size_t)(cnt)))) rc = skg_hip_fail("upload", rc)
    PUT_D(Q->cd, tc, n); PUT_D(Q->ld, tl, n); PUT_D(Q->ud, tu, n); PUT_D(Q->rld, trl, m); PUT_D(Q->rud, tru, m);
#undef PUT_D
    if (!rc && E->mixed) {
        const double *src[5] = {tc, tl, tu, trl, tru};
        const float *dst[5] = {Q->cf, Q->lf, Q->uf, Q->rlf, Q->ruf};
        const int cnt[5] = {n, n, n, m, m};
        float *tmp = (float *)malloc(sizeof(float) * (size_t)((m > n ? m : n) + 1));
        int v;
        if (!tmp) { rc = skg_fail("out of host memory"); goto done; }
        for (v = 0; v < 5 && !rc; ++v) {
            for (k = 0; k < cnt[v]; ++k) tmp[k] = (float)src[v][k];
            if ((rc = skk_h2d((void *)dst[v], tmp, sizeof(float) * (size_t)cnt[v]))) rc = skg_hip_fail("upload", rc);
        }
        free(tmp);
    }
done:
    free(tc); free(tl); free(tu); free(trl); free(tru);
    return rc;
}

/* Row tiles of every product with entry cap `tile` and streaming matrix
 * loads `stream` (A and A^T; a Hessian keeps its own streaming rule); grid
 * cap `grid` for the A^T product and `probe` for the step probe (A, with Q
 * for a QP); an LP's step probe keeps the last `xcache` bytes of x (its
 * longest columns) in LDS (0: none); and the launch shapes that follow. */
static int apply_tiling(engine *E, int tile, int grid, int probe, int xcache, int stream)
{
    skk_pdhg *Q = &E->P;
    if (dcsr_retile(&E->A, E->rp_host[0], tile, probe, stream) ||
        dcsr_retile(&E->AT, E->rp_host[1], tile, grid, stream))
        return -1;
    Q->A = E->A.k;
    Q->AT = E->AT.k;
    Q->probe_grid = E->A.k.grid;
    if (E->has_q) {
        if (dcsr_retile(&E->Q, E->rp_host[2], tile, probe, E->Q.k.stream)) return -1;
        Q->Q = E->Q.k;
        {
            const long long both = (long long)E->A.k.ntiles + E->Q.k.ntiles;
            Q->probe_grid = both < probe ? (int)both : probe;
        }
    }
    Q->fuse_decision = !E->unfused && E->AT.k.grid + Q->probe_grid <= FUSE_MAX_PARTIALS;
    Q->xcache_bytes = E->has_q ? 0 : xcache;
    E->tile_nnz = tile;
    E->grid_cap = grid;
    E->probe_cap = probe;
    E->xcache = Q->xcache_bytes;
    E->stream = stream;
    return 0;
}

/* Build the scaled, permuted, binned problem on the device. */
static int engine_build(engine *E, int m, int n, const int *rp, const int *ci, const double *v,
                        const int *qrp, const int *qci, const double *qv,
                        const sk_gpu_pdlp_params *P)
{
    hcsr A0, AT0, Ap, ATp;
    int *inv_r = NULL, *inv_c = NULL, i, j, k, q, rc = -1;
    skk_bins bA, bAT;
    double tp[6];    /* phase clock: transpose, scaling, order, copies, upload */
    tp[0] = skg_now();
    memset(&Ap, 0, sizeof(Ap)); memset(&ATp, 0, sizeof(ATp)); memset(&AT0, 0, sizeof(AT0));
    E->m = m; E->n = n;
    A0.rows = m; A0.cols = n; A0.nnz = rp[m];
    A0.p = (int *)rp; A0.j = (int *)ci; A0.v = (double *)v;
    if (hcsr_transpose(&A0, &AT0)) return skg_fail("out of host memory");
    tp[1] = skg_now();
    E->R = (double *)malloc(sizeof(double) * (size_t)(m + 1));
    E->C = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    E->perm_r = (int *)malloc(sizeof(int) * (size_t)(m + 1));
    E->perm_c = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    inv_r = (int *)malloc(sizeof(int) * (size_t)(m + 1));
    inv_c = (int *)malloc(sizeof(int) * (size_t)(n + 1));
    if (!E->R || !E->C || !E->perm_r || !E->perm_c || !inv_r || !inv_c) {
        skg_fail("out of host memory"); goto done;
    }
    if (P->scaling) {
        skg_ruiz(&A0, &AT0, P->ruiz_iterations, P->ruiz_tol, E->R, E->C);
        skg_pock_chambolle_into(&A0, &AT0, E->R, E->C);
    } else {
        for (i = 0; i < m; ++i) E->R[i] = 1.0;
        for (j = 0; j < n; ++j) E->C[j] = 1.0;
    }
    tp[2] = skg_now();
    /* Bin classes decide the permutations.  Inside a class, rows are grouped
       by the longest column they hold (ties: the lower index), so that the
       rows of each long column of a power-law matrix sit together and A^T y
       gathers them almost sequentially; columns are then ordered inside a
       class by the first slot of their rows, so the many short columns are
       gathered almost sequentially by A x.  (Measured on the 1M x 2M LP:
       3-4% more iterations per second than rows in their original order,
       the 500k QP 3.6%.) */
    {
        int *first = (int *)malloc(sizeof(int) * (size_t)(n + 1));
        int *rowkey = (int *)malloc(sizeof(int) * (size_t)(m + 1));
        if (!first || !rowkey) { free(first); free(rowkey); skg_fail("out of host memory"); goto done; }
#pragma omp parallel for schedule(static)
        for (i = 0; i < m; ++i) {
            int best = n, bdeg = -1, q2;
            for (q2 = A0.p[i]; q2 < A0.p[i + 1]; ++q2) {
                const int col = A0.j[q2], deg = AT0.p[col + 1] - AT0.p[col];
                if (deg > bdeg || (deg == bdeg && col < best)) { bdeg = deg; best = col; }
            }
            rowkey[i] = best;
        }
        rc = skg_plan_bins_keyed(A0.p, m, &g_skg_tuning, rowkey, n, E->perm_r, &bA);
        if (!rc) {
            for (k = 0; k < m; ++k) inv_r[E->perm_r[k]] = k;
#pragma omp parallel for schedule(static)
            for (j = 0; j < n; ++j) {
                int mn = m, q2;
                for (q2 = AT0.p[j]; q2 < AT0.p[j + 1]; ++q2) if (inv_r[AT0.j[q2]] < mn) mn = inv_r[AT0.j[q2]];
                first[j] = mn;
            }
            rc = skg_plan_bins_keyed(AT0.p, n, &g_skg_tuning, first, m, E->perm_c, &bAT);
        }
        free(first);
        free(rowkey);
        if (rc) { rc = skg_fail("out of host memory"); goto done; }
    }
    for (k = 0; k < m; ++k) inv_r[E->perm_r[k]] = k;
    for (k = 0; k < n; ++k) inv_c[E->perm_c[k]] = k;
    tp[3] = skg_now();
    if (hcsr_alloc(&Ap, m, n, A0.nnz) || hcsr_alloc(&ATp, n, m, A0.nnz)) { skg_fail("out of host memory"); goto done; }
    /* the scaled matrix and its transpose in slot order: offsets first,
       then every row filled independently */
    q = 0;
    for (k = 0; k < m; ++k) {
        const int r = E->perm_r[k];
        Ap.p[k] = q;
        q += A0.p[r + 1] - A0.p[r];
    }
    Ap.p[m] = q;
#pragma omp parallel for schedule(dynamic, 4096) private(i)
    for (k = 0; k < m; ++k) {
        const int r = E->perm_r[k];
        int o = Ap.p[k];
        for (i = A0.p[r]; i < A0.p[r + 1]; ++i, ++o) {
            Ap.j[o] = inv_c[A0.j[i]];
            Ap.v[o] = A0.v[i] * E->R[r] * E->C[A0.j[i]];
        }
    }
    q = 0;
    for (k = 0; k < n; ++k) {
        const int col = E->perm_c[k];
        ATp.p[k] = q;
        q += AT0.p[col + 1] - AT0.p[col];
    }
    ATp.p[n] = q;
#pragma omp parallel for schedule(dynamic, 4096) private(i)
    for (k = 0; k < n; ++k) {
        const int col = E->perm_c[k];
        int o = ATp.p[k];
        for (i = AT0.p[col]; i < AT0.p[col + 1]; ++i, ++o) {
            ATp.j[o] = inv_r[AT0.j[i]];
            ATp.v[o] = AT0.v[i] * E->R[AT0.j[i]] * E->C[col];
        }
    }
    tp[4] = skg_now();
    if ((rc = dcsr_upload(&E->A, &Ap, E->mixed)) || (rc = dcsr_upload(&E->AT, &ATp, E->mixed))) goto done;
    tp[5] = skg_now();
    if (P->verbosity > 0)
        fprintf(stderr, "set-up: transpose %.3f s, scaling %.3f s, ordering %.3f s, permuted copies %.3f s, "
                "upload %.3f s\n", tp[1] - tp[0], tp[2] - tp[1], tp[3] - tp[2], tp[4] - tp[3], tp[5] - tp[4]);
    E->tile_nnz = skg_tile_nnz(Ap.nnz);
    E->grid_cap = E->probe_cap = SKK_SPMV_GRID;
    E->xcache = 0;
    E->stream = E->A.k.stream;
    E->unfused = P->unfused;
    if (P->autotune || P->tile_nnz || P->spmv_grid || P->probe_grid || P->xcache_bytes || P->stream_loads) {
        E->rp_host[0] = (int *)malloc(sizeof(int) * ((size_t)m + 1));
        E->rp_host[1] = (int *)malloc(sizeof(int) * ((size_t)n + 1));
        if (!E->rp_host[0] || !E->rp_host[1]) { rc = skg_fail("out of host memory"); goto done; }
        memcpy(E->rp_host[0], Ap.p, sizeof(int) * ((size_t)m + 1));
        memcpy(E->rp_host[1], ATp.p, sizeof(int) * ((size_t)n + 1));
    }
    if (E->has_q) {
        /* Q' = C Q C in slot order on both sides */
        hcsr Qp;
        if (hcsr_alloc(&Qp, n, n, qrp[n])) { skg_fail("out of host memory"); goto done; }
        q = 0;
        for (k = 0; k < n; ++k) {
            const int r = E->perm_c[k];
            Qp.p[k] = q;
            for (i = qrp[r]; i < qrp[r + 1]; ++i, ++q) {
                Qp.j[q] = inv_c[qci[i]];
                Qp.v[q] = qv[i] * E->C[r] * E->C[qci[i]];
            }
        }
        Qp.p[n] = q;
        rc = dcsr_upload(&E->Q, &Qp, E->mixed);
        if (!rc && E->rp_host[0]) {
            E->rp_host[2] = (int *)malloc(sizeof(int) * ((size_t)n + 1));
            if (!E->rp_host[2]) rc = skg_fail("out of host memory");
            else memcpy(E->rp_host[2], Qp.p, sizeof(int) * ((size_t)n + 1));
        }
        hcsr_free(&Qp);
        if (rc) goto done;
    }
    {
        skk_pdhg *Q = &E->P;
        double *cd, *ld, *ud, *rld, *rud;
        float *cf = NULL, *lf = NULL, *uf = NULL, *rlf = NULL, *ruf = NULL;
        const size_t nn = (size_t)(n ? n : 1), mm = (size_t)(m ? m : 1);
        int cap, eg;
        memset(Q, 0, sizeof(*Q));
        Q->m = m; Q->n = n;
        Q->A = E->A.k; Q->AT = E->AT.k;
        Q->has_q = E->has_q;
        if (E->has_q) {
            Q->Q = E->Q.k;
            if (E->Q.k.ntiles > 0 && (rc = dalloc(E, (void **)&Q->qx_d, 8 * nn))) goto done;
        }
        if ((rc = dalloc(E, (void **)&cd, 8 * nn)) || (rc = dalloc(E, (void **)&ld, 8 * nn)) ||
            (rc = dalloc(E, (void **)&ud, 8 * nn)) || (rc = dalloc(E, (void **)&rld, 8 * mm)) ||
            (rc = dalloc(E, (void **)&rud, 8 * mm))) goto done;
        if (E->mixed) {
            if ((rc = dalloc(E, (void **)&cf, 4 * nn)) || (rc = dalloc(E, (void **)&lf, 4 * nn)) ||
                (rc = dalloc(E, (void **)&uf, 4 * nn)) || (rc = dalloc(E, (void **)&rlf, 4 * mm)) ||
                (rc = dalloc(E, (void **)&ruf, 4 * mm))) goto done;
        }
        Q->cd = cd; Q->ld = ld; Q->ud = ud; Q->rld = rld; Q->rud = rud;
        Q->cf = cf; Q->lf = lf; Q->uf = uf; Q->rlf = rlf; Q->ruf = ruf;
        if ((rc = alloc_iter(E, &Q->it[SKK_F64], sizeof(double), 0))) goto done;
        /* A QP's FP32 iterates are averaged in FP64: measured on three seeds of
           the config-2 generator, 3-12% less time to 1e-6; on LPs (config 0,
           the config-4 family) FP32 sums were 4-10% faster, and they keep
           config 1's iterations lighter. */
        if (E->mixed && (rc = alloc_iter(E, &Q->it[SKK_F32], sizeof(float), E->has_q))) goto done;
        if (E->mixed && P->refine) {
            Q->anchored = 1;
            if ((rc = dalloc(E, (void **)&Q->hx, 8 * nn)) || (rc = dalloc(E, (void **)&Q->hy, 8 * mm)) ||
                (rc = dalloc(E, (void **)&Q->hax, 8 * mm)) || (rc = dalloc(E, (void **)&Q->hg, 8 * nn)) ||
                (rc = dalloc(E, (void **)&Q->ycf, 4 * mm)) ||
                (E->has_q && (rc = dalloc(E, (void **)&Q->hqx, 8 * nn))))
                goto done;
        }
        eg = ((m > n ? m : n) + SKK_BLOCK - 1) / SKK_BLOCK;
        if (eg < 1) eg = 1;
        if (eg > 2048) eg = 2048;
        Q->elem_grid = eg;
        /* a QP's step probe runs Q x and A x in one launch over both tile sets */
        Q->probe_grid = E->A.k.grid;
        if (E->has_q) {
            const long long both = (long long)E->A.k.ntiles + E->Q.k.ntiles;
            Q->probe_grid = both < SKK_SPMV_GRID ? (int)both : SKK_SPMV_GRID;
        }
        cap = eg;
        if (E->A.k.grid > cap) cap = E->A.k.grid;
        if (E->AT.k.grid > cap) cap = E->AT.k.grid;
        if (E->has_q && E->Q.k.grid > cap) cap = E->Q.k.grid;
        if (Q->probe_grid > cap) cap = Q->probe_grid;
        if (E->rp_host[0] && cap < SKK_SPMV_GRID) cap = SKK_SPMV_GRID;   /* any later tiling fits */
        Q->part_capacity = cap;
        /* with few block partials every dual-prox workgroup can afford to
           reduce them itself: the step decision then rides in that launch */
        Q->fuse_decision = !P->unfused && E->AT.k.grid + Q->probe_grid <= FUSE_MAX_PARTIALS;
        if ((rc = dalloc(E, (void **)&Q->ax_d, 8 * mm)) || (rc = dalloc(E, (void **)&Q->aty_d, 8 * nn)) ||
            (rc = dalloc(E, (void **)&Q->st, sizeof(skk_state))) ||
            (rc = dalloc(E, (void **)&Q->snap, sizeof(skk_state))) ||
            (rc = dalloc(E, (void **)&Q->part, sizeof(double) * 16 * (size_t)cap)) ||
            (rc = dalloc(E, (void **)&Q->eval_out, sizeof(double) * 24)) ||
            (rc = dalloc(E, (void **)&E->pv, 8 * nn)) || (rc = dalloc(E, (void **)&E->pw, 8 * mm)) ||
            (rc = dalloc(E, (void **)&E->pv2, 8 * nn)) || (rc = dalloc(E, (void **)&E->pnorm, 8)))
            goto done;
        {
            /* the result leaves the device in the caller's order, unscaled */
            int *cs, *rs;
            double *cf2, *rf2;
            if ((rc = dalloc(E, (void **)&cs, 4 * nn)) || (rc = dalloc(E, (void **)&rs, 4 * mm)) ||
                (rc = dalloc(E, (void **)&cf2, 8 * nn)) || (rc = dalloc(E, (void **)&rf2, 8 * mm)))
                goto done;
            if ((rc = skk_h2d(cs, inv_c, sizeof(int) * (size_t)n)) ||
                (rc = skk_h2d(rs, inv_r, sizeof(int) * (size_t)m)) ||
                (rc = skk_h2d(cf2, E->C, sizeof(double) * (size_t)n)) ||
                (rc = skk_h2d(rf2, E->R, sizeof(double) * (size_t)m))) {
                rc = skg_hip_fail("upload", rc); goto done;
            }
            E->U.col_slot = cs; E->U.row_slot = rs; E->U.col_scale = cf2; E->U.row_scale = rf2;
        }
        if (P->restart_rule == SK_GPU_RESTART_NDG) {
            Q->ndg = 1;
            if ((rc = dalloc(E, (void **)&Q->nx, 8 * nn)) || (rc = dalloc(E, (void **)&Q->ng, 8 * nn)) ||
                (rc = dalloc(E, (void **)&Q->ny, 8 * mm)) || (rc = dalloc(E, (void **)&Q->nax, 8 * mm)) ||
                (rc = dalloc(E, (void **)&Q->ndg_s, sizeof(double) * 8)) ||
                (rc = dalloc(E, (void **)&Q->ndg_part, sizeof(double) * 32 * (size_t)cap)))
                goto done;
        }
        if ((rc = skk_host_alloc((void **)&E->h_eval, sizeof(double) * 24 + sizeof(skk_state)))) {
            rc = skg_hip_fail("hipHostMalloc", rc); goto done;
        }
        E->h_state = (skk_state *)(E->h_eval + 24);
    }
    if (!P->autotune && E->rp_host[0]) {
        /* a tiling fixed by the caller (a tuned choice reused, say) */
        const int tile = P->tile_nnz ? P->tile_nnz : E->tile_nnz;
        const int grid = P->spmv_grid ? P->spmv_grid : SKK_SPMV_GRID;
        const int probe = P->probe_grid ? P->probe_grid : grid;
        const int stream = P->stream_loads ? P->stream_loads == 2 : E->stream;
        if (tile < SKK_TILE_MIN_NNZ || tile > SKK_TILE_NNZ || grid < 1 || grid > SKK_SPMV_GRID ||
            probe < 1 || probe > SKK_SPMV_GRID || P->xcache_bytes < 0 || P->xcache_bytes > SKK_XCACHE_MAX ||
            P->xcache_bytes % 16 ||
            P->stream_loads < 0 || P->stream_loads > 2) {
            rc = skg_fail("invalid tiling: tile_nnz %d, spmv_grid %d, probe_grid %d, xcache_bytes %d, "
                          "stream_loads %d", P->tile_nnz, P->spmv_grid, P->probe_grid, P->xcache_bytes,
                          P->stream_loads);
            goto done;
        }
        if ((rc = apply_tiling(E, tile, grid, probe, P->xcache_bytes, stream))) goto done;
    }
    rc = 0;
done:
    hcsr_free(&AT0); hcsr_free(&Ap); hcsr_free(&ATp);
    free(inv_r); free(inv_c);
    return rc;
}

/* ------------------------------------------------------------------ loop */

typedef struct kkt {
    double primal, dual, gap, pobj, dobj, pmove, dmove;
} kkt;

static double kkt_worst(const kkt *k)
{
    const double a = k->primal > k->dual ? k->primal : k->dual;
    return a > k->gap ? a : k->gap;
}

/* GpuPdlp.cpp:226-249 */
static kkt kkt_from(const double *s, double cost_norm, double bound_norm)
{
    kkt k;
    k.primal = sqrt(fmax(0.0, s[0])) / (1.0 + bound_norm);
    k.dual = sqrt(fmax(0.0, s[1])) / (1.0 + cost_norm);
    k.pobj = s[2]; k.dobj = s[3];
    k.gap = fabs(s[2] - s[3]) / (1.0 + fabs(s[2]) + fabs(s[3]));
    k.pmove = sqrt(fmax(0.0, s[6]));
    k.dmove = sqrt(fmax(0.0, s[7]));
    if (!isfinite(k.primal) || !isfinite(k.dual) || !isfinite(k.gap))
        k.primal = k.dual = k.gap = INFINITY;
    return k;
}

/* Queue one restart window: `period` iterations, the pending-sum flush and
 * both evaluations.  Launch count is returned through E->launches. */
static int queue_window(engine *E, int prec, int period, int adaptive, double fixed_terms)
{
    int k, rc;
    if (adaptive) {
        if ((rc = skk_pdhg_adaptive_run(&E->P, prec, period))) return skg_hip_fail("PDHG iteration launch", rc);
        E->launches += (long long)period * (E->P.fuse_decision ? 3 : 4) + (E->P.fuse_decision ? 1 : 0);
    } else {
        for (k = 0; k < period; ++k)
            if ((rc = skk_pdhg_fixed_iter(&E->P, prec))) return skg_hip_fail("PDHG iteration launch", rc);
        E->launches += (long long)period * 2;
    }
    if (adaptive && (rc = skk_pdhg_flush(&E->P, prec))) return skg_hip_fail("flush launch", rc);
    if ((rc = skk_pdhg_eval(&E->P, prec, 0, adaptive, fixed_terms))) return skg_hip_fail("evaluation launch", rc);
    if ((rc = skk_pdhg_eval(&E->P, prec, 1, adaptive, fixed_terms))) return skg_hip_fail("evaluation launch", rc);
    E->launches += (adaptive ? 2 : 0) + 3 + (adaptive ? 1 : 0) + 4;
    return 0;
}

static int run_window(engine *E, int prec, int period, int adaptive, double fixed_terms, int use_graph,
                      int restart_period)
{
    int rc;
    const int gi = prec;
    if (use_graph && adaptive && period == restart_period) {
        if (!E->graph[gi] || E->graph_period[gi] != period) {
            skk_graph_destroy(E->graph[gi]);
            E->graph[gi] = NULL;
            if ((rc = skk_graph_begin())) return skg_hip_fail("graph capture", rc);
            rc = queue_window(E, prec, period, adaptive, fixed_terms);
            {
                void *exec = NULL;
                const int rc2 = skk_graph_end(&exec);
                if (rc) { skk_graph_destroy(exec); return rc; }
                if (rc2) return skg_hip_fail("graph instantiate", rc2);
                E->graph[gi] = exec;
                E->graph_period[gi] = period;
            }
        }
        if ((rc = skk_graph_launch(E->graph[gi]))) return skg_hip_fail("graph launch", rc);
        E->launches += 1;
    } else {
        if ((rc = queue_window(E, prec, period, adaptive, fixed_terms))) return rc;
    }
    if ((rc = skk_d2h_async(E->h_eval, E->P.eval_out, sizeof(double) * 18)) ||
        (rc = skk_d2h_async(E->h_state, E->P.st, sizeof(skk_state))) || (rc = skk_sync()))
        return skg_hip_fail("window read-back", rc);
    return 0;
}

/* Host state of the restarted loop, kept in the session so that a solve
 * can be resumed with tighter limits from exactly where it stopped (same
 * iterate, step size, primal weight, running averages and restart state). */
typedef struct loop_state {
    int active;                  /* a solve ran and can be resumed          */
    int prec, tail, stall, use_avg_final;
    int iteration, windows_since_restart, its_since_restart;
    double omega, mu_at_restart, mu_previous, best_mu_seen, loop_seconds;
    kkt best;
} loop_state;

struct sk_gpu_pdlp_session {
    engine E;
    sk_gpu_pdlp_params P;
    double norm, eta_global;
    int ready;
    loop_state L;
    sk_gpu_pdlp_stats stats;     /* cumulative over a solve and its resumes */
};

static int check_problem(int rows, int cols, int nnz, const int *rp, const int *ci, const double *v,
                         const double *c, const double *rl, const double *ru, const double *l,
                         const double *u)
{
    int k;
    if (!c || !rl || !ru || !l || !u) return skg_fail("null model pointer");
    if (hcsr_validate(rows, cols, nnz, rp, ci, v)) return -1;
    for (k = 0; k < cols; ++k)
        if (!isfinite(c[k]) || isnan(l[k]) || isnan(u[k]) || l[k] > u[k] ||
            to_ieee(l[k]) == INFINITY || to_ieee(u[k]) == -INFINITY)
            return skg_fail("invalid cost or column bounds at %d", k);
    for (k = 0; k < rows; ++k)
        if (isnan(rl[k]) || isnan(ru[k]) || rl[k] > ru[k] || to_ieee(rl[k]) == INFINITY ||
            to_ieee(ru[k]) == -INFINITY)
            return skg_fail("invalid row bounds at %d", k);
    return 0;
}

static int check_params(const sk_gpu_pdlp_params *P)
{
    if (P->eps_optimal <= 0.0 || P->max_iterations <= 0 || P->restart_period <= 0 ||
        P->power_iterations < 0 || P->eta_min_factor <= 0.0 || P->eta_max_factor < P->eta_min_factor ||
        (P->precision != SK_GPU_PREC_FP64 && P->precision != SK_GPU_PREC_MIXED) ||
        (P->restart_rule != SK_GPU_RESTART_KKT && P->restart_rule != SK_GPU_RESTART_NDG))
        return skg_fail("invalid PDLP parameters");
    return 0;
}

static int session_start(sk_gpu_pdlp_session *S, int prec, const double *x0, const double *y0,
                         double *omega_out);

/* One autotuner candidate: a cold start and a timed run of 32 adaptive
 * iterations (after 8 untimed) on the library stream; milliseconds, or a
 * negative value on a HIP error. */
static double tune_probe(sk_gpu_pdlp_session *S, int prec, int tile, int grid, int probe, int xcache, int stream)
{
    engine *E = &S->E;
    double omega, ms;
    int rc;
    if (apply_tiling(E, tile, grid, probe, xcache, stream) || session_start(S, prec, NULL, NULL, &omega))
        return -1.0;
    if ((rc = S->P.adaptive_step ? skk_pdhg_adaptive_run(&E->P, prec, 8) : skk_pdhg_fixed_iter(&E->P, prec)) ||
        (rc = skk_timeline_mark(0)) ||
        (rc = S->P.adaptive_step ? skk_pdhg_adaptive_run(&E->P, prec, 32) : skk_pdhg_fixed_iter(&E->P, prec)) ||
        (rc = skk_timeline_mark(1)) || (rc = skk_sync())) {
        skg_hip_fail("autotune probe", rc);
        return -1.0;
    }
    ms = skk_timeline_ms(0, 1);
    if (S->P.verbosity > 0)
        fprintf(stderr, "autotune tile %d grid %d probe %d xcache %d stream %d: %.4f ms\n", tile, grid, probe,
                xcache, stream, ms);
    return ms > 0.0 ? ms : INFINITY;
}

/* The autotuner: every tile cap x grid cap x streaming choice with one grid
 * cap for both products and no LDS copy of x; then, for an LP, the best of
 * those with the step probe's x gathered partly from an LDS copy of its
 * longest columns (32 or 64 KB per workgroup) on 512 or 1,024 workgroups
 * (a power-law matrix gathers most of x from a few thousand columns).  The
 * fastest is kept. */
static int autotune(sk_gpu_pdlp_session *S)
{
    static const int tiles[4] = {256, 512, 1024, 2048};
    static const int grids[2] = {1024, SKK_SPMV_GRID};
    static const int caches[2] = {SKK_XCACHE_MAX / 2, SKK_XCACHE_MAX};
    static const int probes[2] = {512, 1024};
    engine *E = &S->E;
    const int prec = E->mixed ? SKK_F32 : SKK_F64;
    const double t0 = skg_now();
    double best_ms = INFINITY, ms;
    int best[5] = {E->tile_nnz, E->grid_cap, E->probe_cap, 0, E->stream}, t, g, st, rc;
    if ((rc = skk_timeline_reserve(2))) return skg_hip_fail("event timeline", rc);
    for (t = 0; t < 4; ++t)
        for (g = 0; g < 2; ++g)
            for (st = 0; st < 2; ++st) {
                if ((ms = tune_probe(S, prec, tiles[t], grids[g], grids[g], 0, st)) < 0.0) return -1;
                if (ms < best_ms) {
                    best_ms = ms;
                    best[0] = tiles[t]; best[1] = best[2] = grids[g]; best[3] = 0; best[4] = st;
                }
            }
    if (!E->has_q) {
        const int t1 = best[0], g1 = best[1], s1 = best[4];
        int c, pg;
        for (c = 0; c < 2; ++c)
            for (pg = 0; pg < 2; ++pg) {
                if ((ms = tune_probe(S, prec, t1, g1, probes[pg], caches[c], s1)) < 0.0) return -1;
                if (ms < best_ms) { best_ms = ms; best[2] = probes[pg]; best[3] = caches[c]; }
            }
    }
    if (apply_tiling(E, best[0], best[1], best[2], best[3], best[4])) return -1;
    E->tuned = 1;
    E->tune_seconds = skg_now() - t0;
    return 0;
}

int sk_gpu_pdlp_session_create(int rows, int cols, int nnz, const int *row_offsets,
                               const int *column_indices, const double *values, const double *c,
                               const double *row_lower, const double *row_upper,
                               const double *col_lower, const double *col_upper,
                               const sk_gpu_pdlp_params *params, sk_gpu_pdlp_session **out)
{
    return sk_gpu_pdlp_session_create_qp(rows, cols, nnz, row_offsets, column_indices, values, NULL,
                                         NULL, NULL, 0, c, row_lower, row_upper, col_lower, col_upper,
                                         params, out);
}

int sk_gpu_pdlp_session_create_qp(int rows, int cols, int nnz, const int *row_offsets,
                                  const int *column_indices, const double *values,
                                  const int *q_row_offsets, const int *q_column_indices,
                                  const double *q_values, int q_nnz, const double *c,
                                  const double *row_lower, const double *row_upper,
                                  const double *col_lower, const double *col_upper,
                                  const sk_gpu_pdlp_params *params, sk_gpu_pdlp_session **out)
{
    sk_gpu_pdlp_session *S;
    double s_norm = 0.0;
    int rc;
    const int has_q = q_row_offsets != NULL;
    skg_clear_error();
    if (!out) return skg_fail("null session pointer");
    *out = NULL;
    if (check_problem(rows, cols, nnz, row_offsets, column_indices, values, c, row_lower, row_upper,
                      col_lower, col_upper)) return -1;
    if (has_q && hcsr_validate(cols, cols, q_nnz, q_row_offsets, q_column_indices, q_values)) return -1;
    if (!has_q && (q_column_indices || q_values || q_nnz)) return skg_fail("Hessian arrays without offsets");
    S = (sk_gpu_pdlp_session *)calloc(1, sizeof(*S));
    if (!S) return skg_fail("out of host memory");
    if (params) S->P = *params; else sk_gpu_pdlp_params_default(&S->P);
    if (check_params(&S->P) || skg_have_device()) { free(S); return -1; }
    if (has_q && !S->P.adaptive_step) { free(S); return skg_fail("a Hessian needs the adaptive step"); }
    if (has_q && S->P.restart_rule == SK_GPU_RESTART_NDG) {
        free(S);
        return skg_fail("NDG restarts are for LPs (the trust-region step of a QP is not a projection)");
    }
    S->E.mixed = S->P.precision == SK_GPU_PREC_MIXED;
    S->E.has_q = has_q;
    if (engine_build(&S->E, rows, cols, row_offsets, column_indices, values, q_row_offsets,
                     q_column_indices, q_values, &S->P) ||
        engine_set_data(&S->E, c, row_lower, row_upper, col_lower, col_upper)) {
        engine_free(&S->E); free(S); return -1;
    }
    /* ||A||_2 of the scaled matrix (GpuPdlp.cpp:179-211) */
    if ((rc = skk_power_norm(&S->E.P, S->E.pv, S->E.pw, S->E.pv2, S->P.power_iterations, S->E.pnorm)) ||
        (rc = skk_d2h(&s_norm, S->E.pnorm, sizeof(double)))) {
        skg_hip_fail("power iteration", rc); engine_free(&S->E); free(S); return -1;
    }
    S->norm = sqrt(sqrt(fmax(0.0, s_norm)));
    if (!(S->norm > 0.0) || !isfinite(S->norm)) S->norm = 1.0;
    S->eta_global = 0.9 / S->norm;
    S->ready = 1;
    if (S->P.autotune && autotune(S)) { engine_free(&S->E); free(S); return -1; }
    *out = S;
    return 0;
}

int sk_gpu_pdlp_session_update(sk_gpu_pdlp_session *S, const double *c, const double *row_lower,
                               const double *row_upper, const double *col_lower,
                               const double *col_upper)
{
    skg_clear_error();
    if (!S || !S->ready) return skg_fail("invalid session");
    if (!c || !row_lower || !row_upper || !col_lower || !col_upper) return skg_fail("null data pointer");
    return engine_set_data(&S->E, c, row_lower, row_upper, col_lower, col_upper) ? -1 : 0;
}

int sk_gpu_pdlp_session_set_limits(sk_gpu_pdlp_session *S, int max_iterations, double time_limit_seconds,
                                   double eps_optimal)
{
    skg_clear_error();
    if (!S || !S->ready) return skg_fail("invalid session");
    if (max_iterations <= 0 || time_limit_seconds < 0.0 || !(eps_optimal > 0.0))
        return skg_fail("invalid limits");
    S->P.max_iterations = max_iterations;
    S->P.time_limit_seconds = time_limit_seconds;
    S->P.eps_optimal = eps_optimal;
    return 0;
}

void sk_gpu_pdlp_session_destroy(sk_gpu_pdlp_session *S)
{
    if (!S) return;
    engine_free(&S->E);
    free(S);
}

/* Cold (x=0, y=0, GpuPdlp.cpp:278-291) or warm start in slot order. */
static int session_start(sk_gpu_pdlp_session *S, int prec, const double *x0, const double *y0,
                         double *omega_out)
{
    engine *E = &S->E;
    const int rows = E->m, cols = E->n;
    double omega = 1.0;
    int k, rc = 0;
    if (S->P.primal_weight > 0.0)
        omega = fmin(1e6, fmax(1e-6, S->P.primal_weight));
    else if (E->cost_norm > 0.0 && E->bound_norm > 0.0)
        omega = fmin(1e6, fmax(1e-6, E->cost_norm / E->bound_norm));
    if (prec == SKK_F32 && E->P.anchored) {
        /* anchored FP32 set: the start point (warm or 0) is the first anchor,
           in FP64; the corrections start at 0 */
        if (x0 || y0) {
            double *tx = (double *)calloc((size_t)cols + 1, sizeof(double));
            double *ty = (double *)calloc((size_t)rows + 1, sizeof(double));
            if (!tx || !ty) { free(tx); free(ty); return skg_fail("out of host memory"); }
            for (k = 0; k < cols && x0; ++k) {
                const int col = E->perm_c[k];
                tx[k] = x0[col] / E->C[col];
            }
            for (k = 0; k < rows && y0; ++k) ty[k] = y0[E->perm_r[k]] / E->R[E->perm_r[k]];
            rc = skk_h2d(E->P.hx, tx, sizeof(double) * (size_t)cols);
            if (!rc) rc = skk_h2d(E->P.hy, ty, sizeof(double) * (size_t)rows);
            free(tx); free(ty);
        } else {
            rc = skk_fill_f64(E->P.hx, 0.0, cols);
            if (!rc) rc = skk_fill_f64(E->P.hy, 0.0, rows);
        }
        if (rc) return skg_hip_fail("start point upload", rc);
        if ((rc = skk_pdhg_init(&E->P, prec, S->eta_global, omega, S->P.eta_min_factor / S->norm,
                                S->P.eta_max_factor / S->norm, 0, 0)) ||
            (rc = skk_pdhg_reanchor(&E->P, 0)))
            return skg_hip_fail("PDHG init", rc);
        *omega_out = omega;
        return 0;
    }
    if (x0 || y0) {
        double *tx = (double *)malloc(sizeof(double) * (size_t)(cols + 1));
        double *ty = (double *)malloc(sizeof(double) * (size_t)(rows + 1));
        float *fx = (float *)malloc(sizeof(float) * (size_t)(cols + 1));
        float *fy = (float *)malloc(sizeof(float) * (size_t)(rows + 1));
        if (!tx || !ty || !fx || !fy) { free(tx); free(ty); free(fx); free(fy); return skg_fail("out of host memory"); }
        for (k = 0; k < cols; ++k) tx[k] = x0 ? x0[E->perm_c[k]] / E->C[E->perm_c[k]] : 0.0;
        for (k = 0; k < rows; ++k) ty[k] = y0 ? y0[E->perm_r[k]] / E->R[E->perm_r[k]] : 0.0;
        if (prec == SKK_F32) {
            for (k = 0; k < cols; ++k) fx[k] = (float)tx[k];
            for (k = 0; k < rows; ++k) fy[k] = (float)ty[k];
            rc = skk_h2d(E->P.it[prec].X[0], fx, sizeof(float) * (size_t)cols);
            if (!rc) rc = skk_h2d(E->P.it[prec].Y[0], fy, sizeof(float) * (size_t)rows);
        } else {
            rc = skk_h2d(E->P.it[prec].X[0], tx, sizeof(double) * (size_t)cols);
            if (!rc) rc = skk_h2d(E->P.it[prec].Y[0], ty, sizeof(double) * (size_t)rows);
        }
        free(tx); free(ty); free(fx); free(fy);
        if (rc) return skg_hip_fail("warm start upload", rc);
    }
    if ((rc = skk_pdhg_init(&E->P, prec, S->eta_global, omega, S->P.eta_min_factor / S->norm,
                            S->P.eta_max_factor / S->norm, x0 != NULL, y0 != NULL)))
        return skg_hip_fail("PDHG init", rc);
    *omega_out = omega;
    return 0;
}

int sk_gpu_pdlp_session_iterate(sk_gpu_pdlp_session *S, int iterations, double *kernel_ms)
{
    engine *E;
    int prec, k, part, rc;
    double omega;
    skg_clear_error();
    if (!S || !S->ready || iterations < 0) return skg_fail("invalid session or count");
    E = &S->E;
    prec = E->mixed ? SKK_F32 : SKK_F64;
    if (session_start(S, prec, NULL, NULL, &omega)) return -1;
    if (kernel_ms) {
        const int parts = S->P.adaptive_step ? 4 : 1;
        if ((rc = skk_timeline_reserve(5 * iterations + 5))) return skg_hip_fail("event timeline", rc);
        if ((rc = skk_sync())) return skg_hip_fail("sync", rc);
        for (k = 0; k < iterations; ++k) {
            if ((rc = skk_timeline_mark(5 * k))) return skg_hip_fail("event record", rc);
            for (part = 0; part < parts; ++part) {
                rc = S->P.adaptive_step ? skk_pdhg_adaptive_part(&E->P, prec, part)
                                        : skk_pdhg_fixed_iter(&E->P, prec);
                if (rc) return skg_hip_fail("PDHG launch", rc);
                if ((rc = skk_timeline_mark(5 * k + part + 1))) return skg_hip_fail("event record", rc);
            }
        }
        if ((rc = skk_sync())) return skg_hip_fail("sync", rc);
        for (part = 0; part < 5; ++part) kernel_ms[part] = 0.0;
        for (k = 0; k < iterations; ++k) {
            for (part = 0; part < parts; ++part)
                kernel_ms[part] += skk_timeline_ms(5 * k + part, 5 * k + part + 1);
            kernel_ms[4] += skk_timeline_ms(5 * k, 5 * k + parts);
        }
        for (part = 0; part < 5; ++part) kernel_ms[part] /= (iterations > 0 ? iterations : 1);
        return 0;
    }
    if (S->P.adaptive_step) {
        /* windows of restart_period iterations, as the solve queues them */
        for (k = 0; k < iterations; k += S->P.restart_period) {
            const int cnt = iterations - k < S->P.restart_period ? iterations - k : S->P.restart_period;
            if ((rc = skk_pdhg_adaptive_run(&E->P, prec, cnt))) return skg_hip_fail("PDHG launch", rc);
        }
    } else {
        for (k = 0; k < iterations; ++k)
            if ((rc = skk_pdhg_fixed_iter(&E->P, prec))) return skg_hip_fail("PDHG launch", rc);
    }
    if ((rc = skk_sync())) return skg_hip_fail("sync", rc);
    return 0;
}

/* The restarted loop (GpuPdlp.cpp:311-499) from the session's loop state,
 * for at most `budget` more iterations, until convergence, the time limit
 * (counted from this call), a cancel or divergence; then the export of the
 * result to x_out, y_out in the original space.  S->stats accumulates. */
static int loop_run(sk_gpu_pdlp_session *S, int budget, double *x_out, double *y_out)
{
    engine *E = &S->E;
    const sk_gpu_pdlp_params *P = &S->P;
    loop_state *L = &S->L;
    sk_gpu_pdlp_stats *stats = &S->stats;
    const double t0 = skg_now();
    const int rows = E->m, cols = E->n;
    const int stop_at = budget > INT_MAX - L->iteration ? INT_MAX : L->iteration + budget;
    int rc, k, diverged = 0;
    stats->converged = 0;
    stats->cancelled = 0;
    while (L->iteration < stop_at) {
        const int period = P->restart_period < stop_at - L->iteration ? P->restart_period : stop_at - L->iteration;
        kkt cur, avg, cand;
        int average_better, sufficient, necessary, artificial;
        double mu;
        if (period <= 0) break;
        if (run_window(E, L->prec, period, P->adaptive_step,
                       (double)(L->windows_since_restart + 1) * (double)P->restart_period, P->use_graph,
                       P->restart_period)) return -1;
        stats->host_syncs++;
        L->iteration += period;
        if (L->prec == SKK_F32) stats->fp32_iterations += period; else stats->fp64_iterations += period;
        L->its_since_restart += period;
        ++L->windows_since_restart;
        cur = kkt_from(E->h_eval, E->cost_norm, E->bound_norm);
        avg = kkt_from(E->h_eval + 8, E->cost_norm, E->bound_norm);
        average_better = kkt_worst(&avg) < kkt_worst(&cur);
        L->best = average_better ? avg : cur;
        stats->restart_merit[0] = E->P.ndg ? E->h_eval[16] : kkt_worst(&cur);
        stats->restart_merit[1] = E->P.ndg ? E->h_eval[17] : kkt_worst(&avg);
        if (P->verbosity > 1)
            fprintf(stderr, "it %d prec %s cur %.3e avg %.3e (p %.2e d %.2e g %.2e) eta %.3e omega %.3e "
                    "merits %.3e %.3e\n", L->iteration, L->prec == SKK_F32 ? "f32" : "f64", kkt_worst(&cur),
                    kkt_worst(&avg), L->best.primal, L->best.dual, L->best.gap, E->h_state->eta,
                    E->h_state->omega, stats->restart_merit[0], stats->restart_merit[1]);
        if (!isfinite(kkt_worst(&cur)) && !isfinite(kkt_worst(&avg))) { diverged = 1; break; }
        {
            static const double marks[4] = {1e-3, 1e-4, 1e-5, 1e-6};
            const double merit = kkt_worst(&L->best);
            for (k = 0; k < 4; ++k)
                if (!stats->iterations_to_merit[k] && merit <= marks[k]) {
                    stats->iterations_to_merit[k] = L->iteration;
                    stats->seconds_to_merit[k] = L->loop_seconds + (skg_now() - t0);
                }
        }
        L->use_avg_final = average_better;
        if (L->best.primal <= P->eps_optimal && L->best.dual <= P->eps_optimal && L->best.gap <= P->eps_optimal) {
            stats->converged = 1;
            break;
        }
        if (P->time_limit_seconds > 0.0 && skg_now() - t0 > P->time_limit_seconds) break;
        /* GpuPdlp.cpp:430-434: cooperative cancel at the window sync */
        if (P->cancel && *P->cancel) { stats->cancelled = 1; break; }
        mu = kkt_worst(&L->best);
        cand = L->best;
        if (E->P.ndg) {
            /* NDG restarts (Applegate et al. 2021, s.3.1): the restart
               candidate is the point of lower normalized duality gap and
               its NDG is the merit the criteria compare; termination stays
               on the relative KKT */
            average_better = stats->restart_merit[1] < stats->restart_merit[0];
            mu = stats->restart_merit[average_better ? 1 : 0];
            cand = average_better ? avg : cur;
        }
        /* mixed precision: hand over to FP64 once FP32 is near its floor.
           Plain FP32 iterates: at tail_switch, or when within 100x of it
           no window has improved the best merit by 10% for
           tail_stall_windows windows.  Anchored iterates have no floor of
           their own (the rounding shrinks with the correction, and every
           restart moves the anchor) and follow the FP64 engine's
           iteration counts (the 500k QP 9,472 vs 9,664, the 20k x 50k
           LP family 14,912 vs 14,912); where they stall, the FP64 engine
           stalls at the same merit (the 1M x 2M LP: relative gap 1.27e-3
           in both), so they never hand over. */
        if (L->prec == SKK_F32 && !L->tail && !E->P.anchored) {
            if (mu < 0.9 * L->best_mu_seen) { L->best_mu_seen = mu; L->stall = 0; }
            else ++L->stall;
            if (mu <= P->tail_switch || (mu <= 100.0 * P->tail_switch && L->stall >= P->tail_stall_windows)) {
                if ((rc = skk_pdhg_restart(&E->P, SKK_F32, average_better, P->adaptive_step)) ||
                    (rc = skk_pdhg_promote(&E->P)) ||
                    (rc = skk_pdhg_eval(&E->P, SKK_F64, 0, P->adaptive_step, 1.0)) ||
                    (rc = skk_pdhg_restart(&E->P, SKK_F64, 0, P->adaptive_step)))
                    return skg_hip_fail("precision switch", rc);
                E->launches += 8;
                L->prec = SKK_F64;
                L->tail = 1;
                L->windows_since_restart = 0;
                L->its_since_restart = 0;
                L->mu_at_restart = mu;
                L->mu_previous = INFINITY;
                L->use_avg_final = 0;
                continue;
            }
        }
        /* restart only when earned (GpuPdlp.cpp:437-478) */
        sufficient = mu <= 0.2 * L->mu_at_restart;
        necessary = mu <= 0.8 * L->mu_at_restart && mu > L->mu_previous;
        artificial = (double)L->its_since_restart >= 0.36 * (double)L->iteration;
        L->mu_previous = mu;
        if (!(sufficient || necessary || artificial)) continue;
        if (P->verbosity > 1)
            fprintf(stderr, "restart at %d: sufficient %d necessary %d artificial %d to %s, merit %.3e\n",
                    L->iteration, sufficient, necessary, artificial, average_better ? "average" : "current", mu);
        if ((rc = skk_pdhg_restart(&E->P, L->prec, average_better, P->adaptive_step)))
            return skg_hip_fail("restart launch", rc);
        E->launches += 1;
        if (L->prec == SKK_F32 && E->P.anchored) {
            /* the restart point becomes the new anchor */
            if ((rc = skk_pdhg_reanchor(&E->P, 1))) return skg_hip_fail("re-anchor", rc);
            E->launches += 5 + (E->has_q ? 2 : 0);
            stats->reanchors++;
        }
        L->use_avg_final = 0;
        L->windows_since_restart = 0;
        L->its_since_restart = 0;
        L->mu_at_restart = mu;
        L->mu_previous = INFINITY;
        stats->restarts++;
        /* primal weight (GpuPdlp.cpp:485-493) */
        if (cand.pmove > 1e-14 && cand.dmove > 1e-14) {
            const double target = cand.dmove / cand.pmove;
            const double updated = exp(0.5 * log(target) + 0.5 * log(L->omega));
            if (P->verbosity > 1) fprintf(stderr, "weight %.4e -> %.4e (target %.4e)\n", L->omega, updated, target);
            if (isfinite(updated)) {
                L->omega = fmin(1e6, fmax(1e-6, updated));
                if ((rc = skk_pdhg_set_omega(&E->P, L->omega))) return skg_hip_fail("primal weight update", rc);
                E->launches += 1;
            }
        }
    }
    if ((rc = skk_pdhg_export(&E->P, L->prec, L->use_avg_final, &E->U, E->pv, E->pw)) ||
        (rc = skk_d2h_async(E->h_state, E->P.st, sizeof(skk_state))) || (rc = skk_sync()))
        return skg_hip_fail("export", rc);
    L->loop_seconds += skg_now() - t0;
    /* the export kernel wrote the caller's order, unscaled (x = D_c x',
       y = D_r y'): the copies land in the output arrays as they are */
    if ((rc = skk_d2h(x_out, E->pv, sizeof(double) * (size_t)cols)) ||
        (rc = skk_d2h(y_out, E->pw, sizeof(double) * (size_t)rows)))
        return skg_hip_fail("result copy", rc);
    stats->host_syncs += 2;
    stats->loop_seconds = L->loop_seconds;
    stats->iterations = L->iteration;
    stats->primal_objective = L->best.pobj;
    stats->dual_objective = L->best.dobj;
    stats->relative_primal_residual = L->best.primal;
    stats->relative_dual_residual = L->best.dual;
    stats->relative_gap = L->best.gap;
    stats->primal_weight = L->omega;
    if (P->adaptive_step) {
        const double rej = E->h_state->updates - E->h_state->accept_count;
        stats->final_step_size = E->h_state->eta;
        stats->step_size_ratio = E->h_state->eta * S->norm;
        stats->rejected_steps = (int)(rej > 0.0 ? rej : 0.0);
    } else {
        stats->final_step_size = S->eta_global;
        stats->step_size_ratio = 0.9;
    }
    stats->kernel_launches = E->launches;
    if (diverged) {
        L->active = 0;
        skg_fail("PDHG iterate became non-finite");
        return -2;
    }
    L->active = 1;
    return 0;
}

int sk_gpu_pdlp_session_solve(sk_gpu_pdlp_session *S, const double *x0, const double *y0,
                              double *x_out, double *y_out, sk_gpu_pdlp_stats *stats)
{
    engine *E;
    const double t0 = skg_now();
    int rc;
    skg_clear_error();
    if (!stats) return skg_fail("null stats");
    memset(stats, 0, sizeof(*stats));
    if (!S || !S->ready) return skg_fail("invalid session");
    E = &S->E;
    if (!x_out || (E->m > 0 && !y_out)) return skg_fail("null output pointer");
    memset(&S->L, 0, sizeof(S->L));
    memset(&S->stats, 0, sizeof(S->stats));
    S->L.best.primal = S->L.best.dual = S->L.best.gap = INFINITY;
    S->L.mu_at_restart = S->L.mu_previous = S->L.best_mu_seen = INFINITY;
    E->launches = 0;
    S->stats.matrix_norm = S->norm;
    S->stats.tile_nnz = E->tile_nnz;
    S->stats.spmv_grid = E->grid_cap;
    S->stats.probe_grid = E->probe_cap;
    S->stats.xcache_bytes = E->xcache;
    S->stats.stream_loads = E->stream;
    S->stats.tuned = E->tuned;
    S->stats.tune_seconds = E->tune_seconds;
    S->L.prec = E->mixed ? SKK_F32 : SKK_F64;
    if (session_start(S, S->L.prec, x0, y0, &S->L.omega)) return -1;
    if ((rc = skk_sync())) return skg_hip_fail("setup", rc);
    S->stats.setup_seconds = skg_now() - t0;
    rc = loop_run(S, S->P.max_iterations, x_out, y_out);
    S->stats.seconds = skg_now() - t0;
    *stats = S->stats;
    return rc;
}

int sk_gpu_pdlp_session_solve_fallback(sk_gpu_pdlp_session *S, const double *x0, const double *y0,
                                       int warm_iterations, double *x_out, double *y_out,
                                       sk_gpu_pdlp_stats *stats, int *used_warm)
{
    const int budget = S ? S->P.max_iterations : 0;
    const double limit = S ? S->P.time_limit_seconds : 0.0;
    const double t0 = skg_now();
    int rc, warm_its = 0;
    skg_clear_error();
    if (!stats || !used_warm) return skg_fail("null stats");
    *used_warm = 0;
    if (!S || !S->ready) return skg_fail("invalid session");
    if (warm_iterations <= 0) return skg_fail("invalid warm iteration budget");
    if (x0 || y0) {
        S->P.max_iterations = warm_iterations < budget ? warm_iterations : budget;
        rc = sk_gpu_pdlp_session_solve(S, x0, y0, x_out, y_out, stats);
        S->P.max_iterations = budget;
        if (rc == -1) return -1;
        if (rc == 0 && stats->converged) { *used_warm = 1; return 0; }
        warm_its = stats->iterations;
        /* the cold start is the fallback, within what is left */
        if (limit > 0.0) {
            S->P.time_limit_seconds = limit - (skg_now() - t0);
            if (S->P.time_limit_seconds <= 0.0) { S->P.time_limit_seconds = limit; return rc; }
        }
    }
    rc = sk_gpu_pdlp_session_solve(S, NULL, NULL, x_out, y_out, stats);
    S->P.time_limit_seconds = limit;
    stats->iterations += warm_its;
    stats->seconds = skg_now() - t0;
    return rc;
}

int sk_gpu_pdlp_session_resume(sk_gpu_pdlp_session *S, double *x_out, double *y_out, sk_gpu_pdlp_stats *stats)
{
    const double t0 = skg_now();
    int rc;
    skg_clear_error();
    if (!stats) return skg_fail("null stats");
    if (!S || !S->ready) return skg_fail("invalid session");
    if (!S->L.active) return skg_fail("no solve to resume on this session");
    if (!x_out || (S->E.m > 0 && !y_out)) return skg_fail("null output pointer");
    rc = loop_run(S, S->P.max_iterations, x_out, y_out);
    S->stats.seconds += skg_now() - t0;
    *stats = S->stats;
    return rc;
}

int sk_gpu_pdlp_solve(int rows, int cols, int nnz, const int *row_offsets, const int *column_indices,
                      const double *values, const int *q_row_offsets, const int *q_column_indices,
                      const double *q_values, int q_nnz, const double *c, const double *row_lower,
                      const double *row_upper, const double *col_lower, const double *col_upper,
                      const double *x0, const double *y0, const sk_gpu_pdlp_params *params,
                      double *x_out, double *y_out, sk_gpu_pdlp_stats *stats)
{
    sk_gpu_pdlp_session *S = NULL;
    const double t0 = skg_now();
    double build;
    int rc;
    skg_clear_error();
    if (!stats) return skg_fail("null stats");
    memset(stats, 0, sizeof(*stats));
    if (!x_out || (rows > 0 && !y_out)) return skg_fail("null output pointer");
    if (sk_gpu_pdlp_session_create_qp(rows, cols, nnz, row_offsets, column_indices, values,
                                      q_row_offsets, q_column_indices, q_values, q_nnz, c, row_lower,
                                      row_upper, col_lower, col_upper, params, &S)) return -1;
    build = skg_now() - t0;
    rc = sk_gpu_pdlp_session_solve(S, x0, y0, x_out, y_out, stats);
    stats->setup_seconds += build;
    stats->seconds = skg_now() - t0;
    stats->host_syncs += 1;
    stats->kernel_launches += 2 + 3LL * (params ? params->power_iterations : 30);
    sk_gpu_pdlp_session_destroy(S);
    return rc;
}
