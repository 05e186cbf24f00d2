/* Host driver of the batched PDHG (C11): per-instance scaling and packing,
 * largest-first work order, workgroup-size autotuning, graph capture of
 * the persistent launch, unscaling. */
#define _POSIX_C_SOURCE 200809L
#include "skg_batch.h"
#include "skg_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void sk_gpu_batch_options_default(sk_gpu_batch_options *o)
{
    memset(o, 0, sizeof(*o));
    sk_gpu_pdlp_params_default(&o->pdlp);
    o->pdlp.max_iterations = 100000;
    o->pdlp.precision = SK_GPU_PREC_MIXED;
    /* small instances: an earlier FP64 handoff measured faster for the whole
       batch (0.79 s against 0.86 s with 32 windows, 4,096 instances) */
    o->pdlp.tail_stall_windows = 8;
    o->block_threads = 0;       /* autotune */
    o->use_graph = 1;
}

static double to_ieee(double b)
{
    if (b >= SK_INFINITY) return INFINITY;
    if (b <= -SK_INFINITY) return -INFINITY;
    return b;
}

static double finite_sq(const double *v, int n)
{
    double a = 0.0;
    int k;
    for (k = 0; k < n; ++k) if (isfinite(v[k])) a += v[k] * v[k];
    return a;
}

typedef struct hbatch {
    int count, rows, cols, nnz, arp, trp;
    skb_inst *inst;
    int *order;
    int *a_rp, *t_rp;
    unsigned short *a_ci, *t_ci;       /* local column indices */
    float *a_v, *t_v;
    double *a_vd, *t_vd;
    double *cd, *ld, *ud, *rld, *rud;
    double *R, *C;             /* scale factors (packed) */
} hbatch;

static void hbatch_free(hbatch *h)
{
    free(h->inst); free(h->order); free(h->a_rp); free(h->a_ci); free(h->t_rp); free(h->t_ci);
    free(h->a_v); free(h->t_v); free(h->a_vd); free(h->t_vd);
    free(h->cd); free(h->ld); free(h->ud); free(h->rld); free(h->rud); free(h->R); free(h->C);
    memset(h, 0, sizeof(*h));
}

static _Thread_local const skb_inst *g_sort_inst;   /* the comparator's batch */
static int by_size_desc(const void *a, const void *b)
{
    const skb_inst *x = &g_sort_inst[*(const int *)a], *y = &g_sort_inst[*(const int *)b];
    const long long sx = (long long)x->m + x->n, sy = (long long)y->m + y->n;
    if (sx != sy) return sx < sy ? 1 : -1;
    return *(const int *)a - *(const int *)b;
}

/* Scale every instance (Ruiz, then Pock-Chambolle) and pack CSR(A),
 * CSR(A^T) and the scaled data. */
static int pack(hbatch *h, int count, const sk_model *models, const sk_gpu_pdlp_params *P)
{
    long long rows = 0, cols = 0, nnz = 0, arp = 0, trp = 0;
    int k;
    memset(h, 0, sizeof(*h));
    h->count = count;
    h->inst = (skb_inst *)calloc((size_t)count + 1, sizeof(skb_inst));
    h->order = (int *)malloc(sizeof(int) * ((size_t)count + 1));
    if (!h->inst || !h->order) return skg_fail("out of host memory");
    for (k = 0; k < count; ++k) {
        const sk_model *md = &models[k];
        skb_inst *d = &h->inst[k];
        d->m = md->nrow; d->n = md->ncol;
        d->row = (int)rows; d->col = (int)cols; d->a_rp = (int)arp; d->t_rp = (int)trp; d->nz = (int)nnz;
        rows += md->nrow; cols += md->ncol; nnz += md->A.p[md->ncol];
        arp += md->nrow + 1; trp += md->ncol + 1;
        h->order[k] = k;
    }
    if (rows > 2147483647LL || cols > 2147483647LL || nnz > 2147483647LL)
        return skg_fail("batch too large for 32-bit offsets");
    h->rows = (int)rows; h->cols = (int)cols; h->nnz = (int)nnz; h->arp = (int)arp; h->trp = (int)trp;
#define AL(p, t, c) if (!((p) = (t *)malloc(sizeof(t) * ((size_t)(c) + 1)))) return skg_fail("out of host memory")
    AL(h->a_rp, int, arp); AL(h->t_rp, int, trp); AL(h->a_ci, unsigned short, nnz); AL(h->t_ci, unsigned short, nnz);
    AL(h->a_v, float, nnz); AL(h->t_v, float, nnz); AL(h->a_vd, double, nnz); AL(h->t_vd, double, nnz);
    AL(h->cd, double, cols); AL(h->ld, double, cols); AL(h->ud, double, cols);
    AL(h->rld, double, rows); AL(h->rud, double, rows); AL(h->R, double, rows); AL(h->C, double, cols);
#undef AL
    /* instances on host threads: each writes its own slices */
    {
        int failed = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(|| : failed)
        for (k = 0; k < count; ++k) {
            const sk_model *md = &models[k];
            skb_inst *d = &h->inst[k];
            hcsr A, AT;
            double *R = h->R + d->row, *Cs = h->C + d->col;
            int i, j, p;
            if (failed) continue;
            if (hcsr_from_csc(&md->A, &A)) { failed = 1; continue; }
            if (hcsr_transpose(&A, &AT)) { hcsr_free(&A); failed = 1; continue; }
            if (P->scaling) {
                skg_ruiz(&A, &AT, P->ruiz_iterations, P->ruiz_tol, R, Cs);
                skg_pock_chambolle_into(&A, &AT, R, Cs);
            } else {
                for (i = 0; i < d->m; ++i) R[i] = 1.0;
                for (j = 0; j < d->n; ++j) Cs[j] = 1.0;
            }
            for (i = 0; i <= d->m; ++i) h->a_rp[d->a_rp + i] = A.p[i];
            for (j = 0; j <= d->n; ++j) h->t_rp[d->t_rp + j] = AT.p[j];
            for (i = 0; i < d->m; ++i)
                for (p = A.p[i]; p < A.p[i + 1]; ++p) {
                    const double v = A.v[p] * R[i] * Cs[A.j[p]];
                    h->a_ci[d->nz + p] = (unsigned short)A.j[p]; h->a_vd[d->nz + p] = v; h->a_v[d->nz + p] = (float)v;
                }
            for (j = 0; j < d->n; ++j)
                for (p = AT.p[j]; p < AT.p[j + 1]; ++p) {
                    const double v = AT.v[p] * R[AT.j[p]] * Cs[j];
                    h->t_ci[d->nz + p] = (unsigned short)AT.j[p]; h->t_vd[d->nz + p] = v; h->t_v[d->nz + p] = (float)v;
                }
            for (j = 0; j < d->n; ++j) {
                h->cd[d->col + j] = md->c[j] * Cs[j];
                h->ld[d->col + j] = to_ieee(md->clow[j]) / Cs[j];
                h->ud[d->col + j] = to_ieee(md->cupp[j]) / Cs[j];
            }
            for (i = 0; i < d->m; ++i) {
                h->rld[d->row + i] = R[i] * to_ieee(md->rlow[i]);
                h->rud[d->row + i] = R[i] * to_ieee(md->rupp[i]);
            }
            d->cost_norm = sqrt(finite_sq(h->cd + d->col, d->n));
            d->bound_norm = sqrt(finite_sq(h->rld + d->row, d->m) + finite_sq(h->rud + d->row, d->m));
            hcsr_free(&A); hcsr_free(&AT);
        }
        if (failed) return skg_fail("out of host memory");
    }
    g_sort_inst = h->inst;
    qsort(h->order, (size_t)count, sizeof(int), by_size_desc);
    return 0;
}

typedef struct dbatch {
    void *p[32];
    int np;
} dbatch;

static int dput(dbatch *d, const void **dst, const void *src, size_t bytes)
{
    void *q = NULL;
    int rc;
    if (d->np >= (int)(sizeof(d->p) / sizeof(d->p[0]))) return skg_fail("too many batch buffers");
    rc = skk_malloc(&q, bytes ? bytes : 8);
    if (rc) return rc;
    d->p[d->np++] = q;
    if (src && bytes && (rc = skk_h2d(q, src, bytes))) return rc;
    *dst = q;
    return 0;
}

static int dput_f(dbatch *d, const float **dst, const double *src, int n)
{
    float *t = (float *)malloc(sizeof(float) * ((size_t)n + 1));
    int k, rc;
    if (!t) return -1;
    for (k = 0; k < n; ++k) t[k] = (float)src[k];
    rc = dput(d, (const void **)dst, t, sizeof(float) * (size_t)n);
    free(t);
    return rc;
}

static void dbatch_free(dbatch *d)
{
    int k;
    for (k = 0; k < d->np; ++k) skk_free(d->p[k]);
    d->np = 0;
}

/* Launch shape of both phases. */
typedef struct phases {
    int bt, grid32, grid64;
    int run32, run64;          /* which phases run */
} phases;

/* The return code of the last launch_phases that failed (0: none). */
static _Thread_local int g_phase_rc;

static int launch_phases(const skb_batch *B, const phases *ph)
{
    int rc = 0;
    if (ph->run32 && (rc = skb_launch(B, ph->bt, ph->grid32, 0))) return g_phase_rc = rc;
    if (ph->run64 && (rc = skb_launch(B, ph->bt, ph->grid64, 1))) return g_phase_rc = rc;
    return 0;
}

/* The error of a failed timed launch, never 0 (a launch refused before it
 * reached the runtime leaves no HIP error behind: -1). */
static int launch_failure(const char *what)
{
    int e = g_phase_rc > 0 ? g_phase_rc : skk_last_launch_error();
    g_phase_rc = 0;
    if (e <= 0) {
        skg_fail("%s refused (block size or grid)", what);
        return -1;
    }
    skg_hip_fail(what, e);
    return e;
}

/* Time the phases (events on the library stream); milliseconds or < 0. */
static double timed_launch(const skb_batch *B, const phases *ph, void **graph, int use_graph)
{
    int rc;
    if (skk_timeline_reserve(2)) return -1.0;
    if (use_graph) {
        if (!*graph) {
            if (skk_graph_begin()) return -1.0;
            rc = launch_phases(B, ph);
            {
                void *exec = NULL;
                const int rc2 = skk_graph_end(&exec);
                if (rc || rc2) { skk_graph_destroy(exec); return -1.0; }
                *graph = exec;
            }
        }
        if (skk_timeline_mark(0)) return -1.0;
        if (skk_graph_launch(*graph)) return -1.0;
    } else {
        if (skk_timeline_mark(0)) return -1.0;
        if (launch_phases(B, ph)) return -1.0;
    }
    if (skk_timeline_mark(1)) return -1.0;
    return skk_timeline_ms(0, 1);
}

#ifndef SKB_MAX_PER_CU
#define SKB_MAX_PER_CU 8
#endif

/* Persistent grid of one phase: as many workgroups as the device holds at
 * once (the runtime's occupancy for this kernel, block size and LDS). */
static int grid_for(int count, int bt, size_t lds_bytes, int fp64)
{
    const int cus = skb_cu_count();
    int per_cu = skb_blocks_per_cu(bt, fp64, lds_bytes);
    if (per_cu <= 0) {
        const int by_lds = lds_bytes > 0 ? (int)(163840 / lds_bytes) : 8;
        per_cu = 2048 / bt;
        if (per_cu > by_lds) per_cu = by_lds;
    }
    if (per_cu > SKB_MAX_PER_CU) per_cu = SKB_MAX_PER_CU;
    if (per_cu < 1) per_cu = 1;
    return count < cus * per_cu ? (count > 0 ? count : 1) : cus * per_cu;
}

int sk_gpu_batch_solve(int count, const sk_model *models, const sk_gpu_batch_options *opt,
                       double *x_out, double *y_out, sk_gpu_batch_result *results,
                       sk_gpu_batch_stats *stats)
{
    sk_gpu_batch_options o;
    hbatch h;
    dbatch d;
    skb_batch B;
    skb_out *hout = NULL;
    double *xs = NULL, *ys = NULL;
    const double t0 = skg_now();
    int k, j, rc = -1, bt, max_m = 0, max_n = 0, mixed;
    size_t lds32, lds64;
    phases ph;
    void *graph = NULL;
    skg_clear_error();
    if (!stats) return skg_fail("null stats");
    memset(stats, 0, sizeof(*stats));
    if (count < 0 || (count > 0 && (!models || !x_out || !y_out || !results)))
        return skg_fail("invalid batch arguments");
    if (opt) o = *opt; else sk_gpu_batch_options_default(&o);
    if (o.pdlp.eps_optimal <= 0.0 || o.pdlp.max_iterations <= 0 || o.pdlp.restart_period <= 0)
        return skg_fail("invalid PDLP parameters");
    for (k = 0; k < count; ++k) {
        if (sk_gpu_model_validate(&models[k]) != SK_OK) return skg_fail("instance %d: invalid model", k);
        if (models[k].Q) return skg_fail("instance %d: quadratic objective not supported in a batch", k);
        for (j = 0; j < models[k].ncol; ++j)
            if (models[k].vartype[j] != SK_CONTINUOUS) return skg_fail("instance %d: integer variable", k);
        if (models[k].nrow > max_m) max_m = models[k].nrow;
        if (models[k].ncol > max_n) max_n = models[k].ncol;
    }
    if (count == 0) return 0;
    if (max_m > 65536 || max_n > 65536)     /* 16-bit local indices (LDS bounds them far lower) */
        return skg_fail("batch instances are limited to 65536 rows and columns");
    if (skg_have_device()) return -1;
    mixed = o.pdlp.precision == SK_GPU_PREC_MIXED;
    if (!mixed && o.pdlp.precision != SK_GPU_PREC_FP64) return skg_fail("unknown precision %d", o.pdlp.precision);
    lds32 = skb_lds_bytes(max_m, max_n, 4);
    lds64 = skb_lds_bytes(max_m, max_n, 8);
    if ((long long)lds32 > (long long)skb_max_lds() || (!mixed && (long long)lds64 > (long long)skb_max_lds()))
        return skg_fail("largest instance needs %zu bytes of LDS (limit %d)", mixed ? lds32 : lds64,
                        skb_max_lds());
    memset(&d, 0, sizeof(d));
    memset(&B, 0, sizeof(B));
    if (pack(&h, count, models, &o.pdlp)) { hbatch_free(&h); return -1; }
    B.count = count; B.max_m = max_m; B.max_n = max_n;
#define PUT(dst, src, bytes) if ((rc = dput(&d, (const void **)&(dst), (src), (bytes)))) goto fail
#define PUTF(dst, src, n) if ((rc = dput_f(&d, &(dst), (src), (n)))) goto fail
    PUT(B.inst, h.inst, sizeof(skb_inst) * (size_t)count);
    PUT(B.order, h.order, sizeof(int) * (size_t)count);
    PUT(B.a_rp, h.a_rp, sizeof(int) * (size_t)h.arp); PUT(B.t_rp, h.t_rp, sizeof(int) * (size_t)h.trp);
    PUT(B.a_ci, h.a_ci, sizeof(unsigned short) * (size_t)h.nnz);
    PUT(B.t_ci, h.t_ci, sizeof(unsigned short) * (size_t)h.nnz);
    PUT(B.a_v, h.a_v, sizeof(float) * (size_t)h.nnz); PUT(B.t_v, h.t_v, sizeof(float) * (size_t)h.nnz);
    PUT(B.a_vd, h.a_vd, sizeof(double) * (size_t)h.nnz); PUT(B.t_vd, h.t_vd, sizeof(double) * (size_t)h.nnz);
    PUT(B.cd, h.cd, sizeof(double) * (size_t)h.cols); PUT(B.ld, h.ld, sizeof(double) * (size_t)h.cols);
    PUT(B.ud, h.ud, sizeof(double) * (size_t)h.cols);
    PUT(B.rld, h.rld, sizeof(double) * (size_t)h.rows); PUT(B.rud, h.rud, sizeof(double) * (size_t)h.rows);
    PUTF(B.c, h.cd, h.cols); PUTF(B.l, h.ld, h.cols); PUTF(B.u, h.ud, h.cols);
    PUTF(B.rl, h.rld, h.rows); PUTF(B.ru, h.rud, h.rows);
    PUT(B.x_out, NULL, sizeof(double) * (size_t)h.cols);
    PUT(B.y_out, NULL, sizeof(double) * (size_t)h.rows);
    PUT(B.out, NULL, sizeof(skb_out) * (size_t)count);
    PUT(B.queue, NULL, sizeof(unsigned));
#undef PUT
#undef PUTF
    B.params.eps = o.pdlp.eps_optimal;
    B.params.eta_min_factor = o.pdlp.eta_min_factor;
    B.params.eta_max_factor = o.pdlp.eta_max_factor;
    B.params.max_iterations = o.pdlp.max_iterations;
    B.params.restart_period = o.pdlp.restart_period;
    B.params.power_iterations = o.pdlp.power_iterations;
    B.params.mixed = mixed;
    B.params.tail_switch = o.pdlp.tail_switch;
    B.params.tail_stall_windows = o.pdlp.tail_stall_windows;
    B.params.anchored = mixed && o.pdlp.refine;
    memset(&ph, 0, sizeof(ph));
    ph.run32 = mixed;
    ph.run64 = 1;
    if (B.params.anchored) {
        /* anchored FP32 iterates: no FP64 tail */
        ph.run64 = 0;
    } else if (mixed && (long long)lds64 > (long long)skb_max_lds()) {
        /* FP64 iterates of the largest instance do not fit in LDS: FP32 only */
        ph.run64 = 0;
        B.params.tail_switch = -1.0;
        B.params.tail_stall_windows = 0x7fffffff;
    }
    stats->setup_seconds = skg_now() - t0;
    /* autotune the workgroup size: time four restart windows of a sample of
       instances spread over the size order, as many as one wave of the
       grid holds -- long enough that the per-iteration rate, not the
       start-up (power iteration, first windows), decides -- with FP32 and
       (when the FP64 tail runs) with FP64 iterates */
    bt = o.block_threads;
    if (bt == 0) {
        const int cand[3] = {128, 256, 512};
        const int sample = count < 4 * skb_cu_count() ? count : 4 * skb_cu_count();
        double best = 1e300;
        skb_batch T = B;
        int *horder = (int *)malloc(sizeof(int) * (size_t)sample);
        if (!horder) { rc = skg_fail("out of host memory"); goto fail; }
        for (k = 0; k < sample; ++k) horder[k] = h.order[(int)(((long long)k * count) / sample)];
        rc = dput(&d, (const void **)&T.order, horder, sizeof(int) * (size_t)sample);
        free(horder);
        if (rc) goto fail;
        T.count = sample;
        T.params.max_iterations = 4 * o.pdlp.restart_period;
        for (k = 0; k < 3; ++k) {
            void *none = NULL;
            phases tp = ph;
            double ms;
            tp.bt = cand[k];
            tp.grid32 = grid_for(sample, cand[k], lds32, 0);
            tp.grid64 = grid_for(sample, cand[k], lds64, 1);
            ms = timed_launch(&T, &tp, &none, 0);
            if (ms < 0.0) { rc = launch_failure("autotune launch"); goto fail; }
            if (mixed && ph.run64) {
                /* four windows hand nothing to the FP64 tail, which is most
                   of a batch's time and whose LDS and occupancy differ:
                   time four windows of FP64 iterates on the sample too */
                skb_batch T64 = T;
                phases t64 = tp;
                double ms64;
                T64.params.mixed = 0;
                t64.run32 = 0;
                t64.run64 = 1;
                ms64 = timed_launch(&T64, &t64, &none, 0);
                if (ms64 < 0.0) { rc = launch_failure("autotune launch"); goto fail; }
                ms += ms64;
            }
            stats->tune_ms[k] = ms;
            if (ms < best) best = ms;
        }
        /* the largest size within 25% of the fastest: the probe is short
           and misses the end of a batch, where a few long instances run
           alone and more lanes shorten their iterations.  4,096 LPs (FP32
           iterates anchored in FP64): the probe puts 512 threads 8-11%
           behind 256 (3.13-3.22 ms against 2.90-2.91 ms), while the whole
           batch takes 0.28 s with 512 and 0.34 s with 256, and a lone slow
           instance 4.5 us per iteration with 512 against 6.5 us with 256 */
        for (k = 0; k < 3; ++k)
            if (stats->tune_ms[k] <= 1.25 * best) bt = cand[k];
    }
    if (bt != 64 && bt != 128 && bt != 256 && bt != 512 && bt != 1024) {
        skg_fail("block_threads must be 64, 128, 256, 512 or 1024"); rc = -1; goto fail;
    }
    ph.bt = bt;
    ph.grid32 = grid_for(count, bt, lds32, 0);
    ph.grid64 = grid_for(count, bt, lds64, 1);
    {
        const double ms = timed_launch(&B, &ph, &graph, o.use_graph);
        if (ms < 0.0) { rc = launch_failure("batch launch"); goto fail; }
        stats->device_seconds = ms * 1e-3;
    }
    hout = (skb_out *)malloc(sizeof(skb_out) * (size_t)count);
    xs = (double *)malloc(sizeof(double) * ((size_t)h.cols + 1));
    ys = (double *)malloc(sizeof(double) * ((size_t)h.rows + 1));
    if (!hout || !xs || !ys) { rc = skg_fail("out of host memory"); goto fail; }
    if ((rc = skk_d2h(hout, B.out, sizeof(skb_out) * (size_t)count)) ||
        (rc = skk_d2h(xs, B.x_out, sizeof(double) * (size_t)h.cols)) ||
        (rc = skk_d2h(ys, B.y_out, sizeof(double) * (size_t)h.rows))) { skg_hip_fail("read-back", rc); goto fail; }
    for (k = 0; k < count; ++k) {
        const skb_inst *di = &h.inst[k];
        const sk_model *md = &models[k];
        double obj = 0.0;
        for (j = 0; j < di->n; ++j) {
            const double xv = xs[di->col + j] * h.C[di->col + j];
            x_out[di->col + j] = xv;
            obj += md->c[j] * xv;
        }
        for (j = 0; j < di->m; ++j) y_out[di->row + j] = ys[di->row + j] * h.R[di->row + j];
        results[k].status = hout[k].converged ? 0 : 1;
        results[k].iterations = hout[k].iterations;
        results[k].restarts = hout[k].restarts;
        results[k].fp32_iterations = hout[k].fp32_iterations;
        results[k].relative_primal_residual = hout[k].rel_primal;
        results[k].relative_dual_residual = hout[k].rel_dual;
        results[k].relative_gap = hout[k].rel_gap;
        results[k].objective = obj + md->objshift;
        stats->converged += hout[k].converged;
        stats->total_iterations += hout[k].iterations;
        stats->fp32_iterations += hout[k].fp32_iterations;
        stats->handoffs += hout[k].fp32_iterations > 0 && hout[k].fp32_iterations < hout[k].iterations;
    }
    stats->block_threads = bt;
    stats->grid = mixed ? ph.grid32 : ph.grid64;
    stats->lds_bytes = (int)(mixed ? lds32 : lds64);
    stats->grid_fp64 = ph.grid64;
    stats->lds_bytes_fp64 = (int)lds64;
    stats->seconds = skg_now() - t0;
    rc = 0;
fail:
    skk_graph_destroy(graph);
    free(hout); free(xs); free(ys);
    dbatch_free(&d);
    hbatch_free(&h);
    if (rc > 0) skg_hip_fail("batch", rc);
    return rc ? -1 : 0;
}
