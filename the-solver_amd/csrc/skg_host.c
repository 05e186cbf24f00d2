/* Host side of sankhya_gpu (C11): model helpers, CSR construction and
 * validation, row-length binning, scaling, the restarted-PDHG (PDLP) control
 * loop, and the public sk_gpu_* C ABI.  Every device operation goes through
 * the thin skk_* layer (skg_kernels.h); this file includes no HIP header. */
#define _POSIX_C_SOURCE 200809L
#include "sankhya_gpu.h"
#include "skg_internal.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>
#include <time.h>

/* ----------------------------------------------------------- errors */

static _Thread_local char g_err[512] = "";

const char *sk_gpu_last_error(void) { return g_err; }

int skg_fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return -1;
}

int skg_hip_fail(const char *what, int code)
{
    snprintf(g_err, sizeof(g_err), "%s: %s", what, skk_error_text(code));
    return code > 0 ? code : -1;
}

void skg_clear_error(void)
{
    g_err[0] = '\0';
    skk_reset_error();
}

double skg_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1.0e-9 * (double)ts.tv_nsec;
}

int skg_have_device(void)
{
    int count = 0;
    const int rc = skk_device_count(&count);
    if (rc) return skg_hip_fail("hipGetDeviceCount", rc);
    if (count == 0) return skg_fail("no HIP device available");
    return 0;
}

/* --------------------------------------------------- model / solution */

void sk_gpu_model_init(sk_model *m)
{
    memset(m, 0, sizeof(*m));
    m->objsense = 1.0;
}

static void free_names(char **v, int n)
{
    int i;
    if (!v) return;
    for (i = 0; i < n; i++) free(v[i]);
    free(v);
}

void sk_gpu_model_free(sk_model *m)
{
    if (!m) return;
    free(m->A.p); free(m->A.i); free(m->A.x);
    if (m->Q) { free(m->Q->p); free(m->Q->i); free(m->Q->x); free(m->Q); }
    free(m->c); free(m->rlow); free(m->rupp); free(m->clow); free(m->cupp); free(m->vartype);
    free_names(m->rowname, m->nrow);
    free_names(m->colname, m->ncol);
    free(m->name);
    sk_gpu_model_init(m);
}

sk_status sk_gpu_model_alloc(sk_model *m, int nrow, int ncol, int nzmax)
{
    if (!m || nrow < 0 || ncol < 0 || nzmax < 0) return SK_ERR_ARG;
    sk_gpu_model_init(m);
    m->nrow = nrow; m->ncol = ncol;
    m->A.nrow = nrow; m->A.ncol = ncol; m->A.nzmax = nzmax;
    m->A.p = (int *)calloc((size_t)ncol + 1, sizeof(int));
    m->A.i = (int *)malloc((size_t)(nzmax > 0 ? nzmax : 1) * sizeof(int));
    m->A.x = (double *)malloc((size_t)(nzmax > 0 ? nzmax : 1) * sizeof(double));
    m->c = (double *)calloc((size_t)ncol + 1, sizeof(double));
    m->clow = (double *)calloc((size_t)ncol + 1, sizeof(double));
    m->cupp = (double *)calloc((size_t)ncol + 1, sizeof(double));
    m->rlow = (double *)calloc((size_t)nrow + 1, sizeof(double));
    m->rupp = (double *)calloc((size_t)nrow + 1, sizeof(double));
    m->vartype = (unsigned char *)calloc((size_t)ncol + 1, 1);
    if (!m->A.p || !m->A.i || !m->A.x || !m->c || !m->clow || !m->cupp || !m->rlow || !m->rupp ||
        !m->vartype) { sk_gpu_model_free(m); return SK_ERR_MEMORY; }
    return SK_OK;
}

static sk_status csc_validate(const sk_csc *a, int rows, int cols)
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

sk_status sk_gpu_model_validate(const sk_model *m)   /* c/src/sk_core.c:140-155 */
{
    int i;
    if (!m || m->nrow < 0 || m->ncol < 0) return SK_ERR_ARG;
    if (csc_validate(&m->A, m->nrow, m->ncol) != SK_OK) return SK_ERR_STRUCTURE;
    if ((m->ncol && (!m->c || !m->clow || !m->cupp || !m->vartype)) ||
        (m->nrow && (!m->rlow || !m->rupp))) return SK_ERR_STRUCTURE;
    for (i = 0; i < m->ncol; ++i)
        if (!isfinite(m->c[i]) || isnan(m->clow[i]) || isnan(m->cupp[i]) || m->clow[i] > m->cupp[i] ||
            (m->vartype[i] != SK_CONTINUOUS && m->vartype[i] != SK_INTEGER)) return SK_ERR_STRUCTURE;
    for (i = 0; i < m->nrow; ++i)
        if (isnan(m->rlow[i]) || isnan(m->rupp[i]) || m->rlow[i] > m->rupp[i]) return SK_ERR_STRUCTURE;
    if (m->Q && csc_validate(m->Q, m->ncol, m->ncol) != SK_OK) return SK_ERR_STRUCTURE;
    return SK_OK;
}

void sk_gpu_solution_init(sk_solution *s)            /* c/src/sk_core.c:166-173 */
{
    memset(s, 0, sizeof(*s));
    s->result = SK_RESULT_NOT_RUN;
    s->objective = NAN;
    s->dual_bound = -SK_INFINITY;
    s->mip_gap = INFINITY;
}

void sk_gpu_solution_free(sk_solution *s)
{
    if (!s) return;
    free(s->x); free(s->y); free(s->rc); free(s->rowact);
    sk_gpu_solution_init(s);
}

void sk_gpu_options_default(sk_options *o)           /* c/src/sk_core.c:184-207 */
{
    memset(o, 0, sizeof(*o));
    o->primal_tol = 1e-7; o->dual_tol = 1e-7; o->pivot_tol = 0.1;
    o->presolve = 1; o->scaling = 1; o->refactor_interval = 50; o->verbosity = 1;
    o->threads = 1; o->mip_gap_abs = 1e-9; o->mip_gap_rel = 1e-6; o->mip_cuts = 1;
    o->mip_heuristics = 1; o->dual_simplex = 1; o->lp_engine = SK_LP_AUTO;
    o->random_seed = 20260825ULL;
}

const char *sk_gpu_result_name(sk_result r)
{
    switch (r) {
    case SK_RESULT_OPTIMAL: return "optimal";
    case SK_RESULT_INFEASIBLE: return "infeasible";
    case SK_RESULT_UNBOUNDED: return "unbounded";
    case SK_RESULT_ITERATION_LIMIT: return "iteration_limit";
    case SK_RESULT_TIME_LIMIT: return "time_limit";
    case SK_RESULT_NUMERIC_FAILURE: return "numeric_failure";
    case SK_RESULT_GAP_LIMIT: return "gap_limit";
    default: return "not_run";
    }
}

const char *sk_gpu_status_name(sk_status s)
{
    switch (s) {
    case SK_OK: return "ok";
    case SK_ERR_ARG: return "invalid_argument";
    case SK_ERR_MEMORY: return "out_of_memory";
    case SK_ERR_STRUCTURE: return "invalid_structure";
    case SK_ERR_IO: return "io_error";
    case SK_ERR_SINGULAR: return "singular";
    case SK_ERR_UNSUPPORTED: return "unsupported";
    case SK_ERR_NUMERIC: return "numeric_error";
    default: return "unknown";
    }
}

/* ----------------------------------------------------------- host CSR */

void hcsr_free(hcsr *a)
{
    if (!a) return;
    free(a->p); free(a->j); free(a->v);
    memset(a, 0, sizeof(*a));
}

int hcsr_alloc(hcsr *a, int rows, int cols, long long nnz)
{
    memset(a, 0, sizeof(*a));
    a->rows = rows; a->cols = cols; a->nnz = (int)nnz;
    a->p = (int *)calloc((size_t)rows + 1, sizeof(int));
    a->j = (int *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
    a->v = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
    if (!a->p || !a->j || !a->v) { hcsr_free(a); return -1; }
    return 0;
}

/* Transpose by host threads with the serial counting sort's exact output:
 * each thread owns a contiguous range of output rows (source columns),
 * scans the whole source in order and places only the entries of its
 * range -- sequential reads, no atomics, and every output row filled in
 * ascending source position.  `cur` holds cols + 1 ints of cursors; the
 * threads use disjoint ranges of it. */
static void transpose_threaded(const hcsr *a, hcsr *t, int *cur)
{
    const int rows = a->rows, cols = a->cols;
#pragma omp parallel
    {
        const int nt = omp_get_num_threads(), id = omp_get_thread_num();
        const int c0 = (int)((long long)cols * id / nt), c1 = (int)((long long)cols * (id + 1) / nt);
        int i, k;
        for (k = 0; k < a->nnz; ++k) {
            const int c = a->j[k];
            if (c >= c0 && c < c1) t->p[c + 1]++;
        }
#pragma omp barrier
#pragma omp single
        for (i = 0; i < cols; ++i) t->p[i + 1] += t->p[i];
        for (i = c0; i < c1; ++i) cur[i] = t->p[i];
        for (i = 0; i < rows; ++i)
            for (k = a->p[i]; k < a->p[i + 1]; ++k) {
                const int c = a->j[k];
                if (c >= c0 && c < c1) {
                    const int q = cur[c]++;
                    t->j[q] = i; t->v[q] = a->v[k];
                }
            }
    }
}

/* Transpose by counting sort; entries of each output row come out in
 * ascending source-row order (source order within a row). */
int hcsr_transpose(const hcsr *a, hcsr *t)
{
    int i, k;
    int *next;
    if (hcsr_alloc(t, a->cols, a->rows, a->nnz)) return -1;
    next = (int *)malloc(((size_t)a->cols + 1) * sizeof(int));
    if (!next) { hcsr_free(t); return -1; }
    if (a->nnz >= (1 << 20) && omp_get_max_threads() > 1) {
        transpose_threaded(a, t, next);
        free(next);
        return 0;
    }
    for (k = 0; k < a->nnz; ++k) t->p[a->j[k] + 1]++;
    for (i = 0; i < a->cols; ++i) t->p[i + 1] += t->p[i];
    memcpy(next, t->p, ((size_t)a->cols + 1) * sizeof(int));
    for (i = 0; i < a->rows; ++i)
        for (k = a->p[i]; k < a->p[i + 1]; ++k) {
            const int q = next[a->j[k]]++;
            t->j[q] = i; t->v[q] = a->v[k];
        }
    free(next);
    return 0;
}

/* CSR view of CSC(A) is CSR(A^T); transposing it gives CSR(A). */
int hcsr_from_csc(const sk_csc *c, hcsr *out)
{
    hcsr at;
    at.rows = c->ncol; at.cols = c->nrow; at.nnz = c->p[c->ncol];
    at.p = c->p; at.j = c->i; at.v = c->x;
    return hcsr_transpose(&at, out);
}

/* Canonical CSR (cuda/src/sankhya_cuda.cu:120-131 rules). */
int hcsr_validate(int rows, int cols, int nnz, const int *rp, const int *ci, const double *v)
{
    int r, p;
    if (rows < 0 || cols < 0 || nnz < 0 || !rp || (nnz > 0 && (!ci || !v)))
        return skg_fail("invalid CSR input");
    if (rp[0] != 0 || rp[rows] != nnz) return skg_fail("invalid CSR row-offset endpoints");
    for (r = 0; r < rows; ++r)
        if (rp[r] < 0 || rp[r] > rp[r + 1] || rp[r + 1] > nnz)
            return skg_fail("invalid CSR row-offset ordering");
    for (p = 0; p < nnz; ++p)
        if (ci[p] < 0 || ci[p] >= cols || !isfinite(v[p]))
            return skg_fail("invalid CSR column index or coefficient");
    return 0;
}

/* ----------------------------------------------------------- binning */

static int ceil_log2(int v)
{
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}

int skg_lanes_log2(int len, const skg_tuning *t)
{
    int need, l;
    if (len > t->long_row) return 8;
    if (len <= t->nnz_per_lane) return 0;
    need = (len + t->nnz_per_lane - 1) / t->nnz_per_lane;
    l = ceil_log2(need);
    return l > 6 ? 6 : l;
}

void skg_tuning_default(skg_tuning *t)
{
    t->nnz_per_lane = 4;
    t->long_row = 512;
}

/* Counting sort of rows by lane class: order[k] = row in slot k.  Inside a
 * class rows keep ascending `key` order when a key is given (a counting sort
 * by key first, key values in [0, key_range]), else their own order.  Fills
 * bins (contiguous slot ranges per class, at least one block). */
int skg_plan_bins_keyed(const int *rp, int rows, const skg_tuning *t, const int *key, int key_range,
                        int *order, skk_bins *bins)
{
    int cls_count[9] = {0}, cls_start[9], k, c, nb = 0, blocks = 0;
    unsigned char *cls = (unsigned char *)malloc((size_t)(rows > 0 ? rows : 1));
    int *pre = NULL;
    if (!cls) return -1;
    for (k = 0; k < rows; ++k) {
        cls[k] = (unsigned char)skg_lanes_log2(rp[k + 1] - rp[k], t);
        cls_count[cls[k]]++;
    }
    cls_start[0] = 0;
    for (c = 1; c < 9; ++c) cls_start[c] = cls_start[c - 1] + cls_count[c - 1];
    memset(bins, 0, sizeof(*bins));
    for (c = 0; c < 9; ++c) {
        int rpb, nblk;
        if (!cls_count[c]) continue;
        rpb = c == 8 ? 1 : (SKK_BLOCK >> c);
        nblk = (cls_count[c] + rpb - 1) / rpb;
        bins->b[nb].row_begin = cls_start[c];
        bins->b[nb].row_end = cls_start[c] + cls_count[c];
        bins->b[nb].log2L = c;
        bins->b[nb].block_begin = blocks;
        blocks += nblk;
        ++nb;
    }
    if (nb == 0) {   /* no rows: one idle block keeps fused epilogues running */
        bins->b[0].row_begin = bins->b[0].row_end = 0;
        bins->b[0].log2L = 0;
        bins->b[0].block_begin = 0;
        nb = 1; blocks = 1;
    }
    bins->nbins = nb;
    bins->total_blocks = blocks;
    if (order) {
        int next[9];
        if (key) {
            /* rows in ascending key order (stable), then a stable pass by class */
            int *cnt = (int *)calloc((size_t)key_range + 2, sizeof(int));
            pre = (int *)malloc(sizeof(int) * (size_t)(rows > 0 ? rows : 1));
            if (!cnt || !pre) { free(cnt); free(pre); free(cls); return -1; }
            for (k = 0; k < rows; ++k) cnt[key[k] + 1]++;
            for (k = 0; k <= key_range; ++k) cnt[k + 1] += cnt[k];
            for (k = 0; k < rows; ++k) pre[cnt[key[k]]++] = k;
            free(cnt);
        }
        memcpy(next, cls_start, sizeof(next));
        for (k = 0; k < rows; ++k) {
            const int r = pre ? pre[k] : k;
            order[next[cls[r]]++] = r;
        }
    }
    free(pre);
    free(cls);
    return 0;
}

int skg_plan_bins(const int *rp, int rows, const skg_tuning *t, int *order, skk_bins *bins)
{
    return skg_plan_bins_keyed(rp, rows, t, NULL, 0, order, bins);
}

/* ----------------------------------------------------------- device CSR */

/* Entry cap of the tiles of a matrix with nnz entries: the largest power
 * of two from SKK_TILE_NNZ down to SKK_TILE_MIN_NNZ that still leaves at
 * least SKK_MIN_TILES tiles, so that a small matrix spreads over enough
 * workgroups to occupy the chip while a large one keeps full tiles. */
int skg_tile_nnz(long long nnz)
{
    int t = SKK_TILE_NNZ;
    while (t > SKK_TILE_MIN_NNZ && nnz / t < SKK_MIN_TILES) t >>= 1;
    return t;
}

/* CSR-stream tiles of consecutive rows (greedy, in row order): a tile
 * closes before it would exceed tile_nnz (<= SKK_TILE_NNZ) entries or
 * SKK_TILE_ROWS rows; a row longer than tile_nnz is a tile of its own.  At
 * least one (possibly empty) tile, so fused epilogues always run.  Returns
 * the tile count, or -1 when out of memory or tile_nnz is out of range;
 * *tile_row receives malloc'd offsets. */
int skg_plan_tiles(const int *rp, int rows, int tile_nnz, int **tile_row)
{
    int cap = rows / 64 + 16, nt = 0, r = 0;
    int *t;
    if (tile_nnz < 1 || tile_nnz > SKK_TILE_NNZ) return -1;
    t = (int *)malloc(sizeof(int) * (size_t)cap);
    if (!t) return -1;
    t[0] = 0;
    while (r < rows || nt == 0) {
        int nnz = 0, cnt = 0;
        if (nt + 2 > cap) {
            int *g;
            cap *= 2;
            g = (int *)realloc(t, sizeof(int) * (size_t)cap);
            if (!g) { free(t); return -1; }
            t = g;
        }
        if (r < rows && rp[r + 1] - rp[r] > tile_nnz) {
            ++r;
        } else {
            while (r < rows && cnt < SKK_TILE_ROWS && nnz + (rp[r + 1] - rp[r]) <= tile_nnz) {
                nnz += rp[r + 1] - rp[r];
                ++r;
                ++cnt;
            }
        }
        t[++nt] = r;
    }
    *tile_row = t;
    return nt;
}

/* Device tile descriptors of the nt tiles whose first rows are tile_row[]:
 * pairs (first row, first entry), nt + 1 of them, so that one load gives a
 * workgroup both ends of its tile's rows and entries. */
static int upload_tiles(const int *rp, const int *tile_row, int nt, int **dev)
{
    int rc, t, *h = (int *)malloc(sizeof(int) * 2 * ((size_t)nt + 1));
    *dev = NULL;
    if (!h) return skg_fail("out of host memory");
    for (t = 0; t <= nt; ++t) {
        h[2 * t] = tile_row[t];
        h[2 * t + 1] = rp[tile_row[t]];
    }
    if ((rc = skk_malloc((void **)dev, sizeof(int) * 2 * ((size_t)nt + 1))) ||
        (rc = skk_h2d(*dev, h, sizeof(int) * 2 * ((size_t)nt + 1)))) {
        free(h); skk_free(*dev); *dev = NULL;
        return skg_hip_fail("tile upload", rc);
    }
    free(h);
    return 0;
}

/* Replace a device CSR's row tiles: entry cap tile_nnz, grid cap grid_cap
 * (<= SKK_SPMV_GRID), non-temporal matrix loads when stream is set.  rp is
 * the host copy of the row offsets the matrix was uploaded with. */
int dcsr_retile(dcsr *d, const int *rp, int tile_nnz, int grid_cap, int stream)
{
    int nt, *tiles = NULL, *dt = NULL;
    if (grid_cap < 1 || grid_cap > SKK_SPMV_GRID) return skg_fail("invalid grid cap %d", grid_cap);
    nt = skg_plan_tiles(rp, d->k.rows, tile_nnz, &tiles);
    if (nt < 0) return skg_fail("invalid tile size %d or out of host memory", tile_nnz);
    if (upload_tiles(rp, tiles, nt, &dt)) { free(tiles); return -1; }
    free(tiles);
    skk_free(d->tiles);
    d->tiles = dt;
    d->k.tiles = dt;
    d->k.ntiles = nt;
    d->k.grid = nt < grid_cap ? nt : grid_cap;
    d->k.stream = stream;
    return 0;
}

void dcsr_free(dcsr *d)
{
    skk_free(d->rp); skk_free(d->ci); skk_free(d->vd); skk_free(d->vf); skk_free(d->tiles);
    memset(d, 0, sizeof(*d));
}

/* Upload a host CSR (rows in the order the products run) with its tiles. */
int dcsr_upload(dcsr *d, const hcsr *a, int with_f32)
{
    int rc, nt, *tiles = NULL;
    memset(d, 0, sizeof(*d));
    d->k.rows = a->rows; d->k.cols = a->cols; d->k.nnz = a->nnz;
    nt = skg_plan_tiles(a->p, a->rows, skg_tile_nnz(a->nnz), &tiles);
    if (nt < 0) return skg_fail("out of host memory");
    d->k.ntiles = nt;
    d->k.grid = nt < SKK_SPMV_GRID ? nt : SKK_SPMV_GRID;
    d->k.stream = (long long)a->nnz * 12 > SKK_STREAM_BYTES;
    if ((rc = skk_malloc((void **)&d->rp, sizeof(int) * ((size_t)a->rows + 1))) ||
        (rc = skk_malloc((void **)&d->ci, sizeof(int) * (size_t)(a->nnz ? a->nnz : 1))) ||
        (rc = skk_malloc((void **)&d->vd, sizeof(double) * (size_t)(a->nnz ? a->nnz : 1))))
        goto fail;
    if ((rc = skk_h2d(d->rp, a->p, sizeof(int) * ((size_t)a->rows + 1))) ||
        (rc = skk_h2d(d->ci, a->j, sizeof(int) * (size_t)a->nnz)) ||
        (rc = skk_h2d(d->vd, a->v, sizeof(double) * (size_t)a->nnz)))
        goto fail;
    if (upload_tiles(a->p, tiles, nt, &d->tiles)) { free(tiles); dcsr_free(d); return -1; }
    if (with_f32) {
        float *tmp = (float *)malloc(sizeof(float) * (size_t)(a->nnz ? a->nnz : 1));
        int k;
        if (!tmp) { rc = -1; goto fail; }
#pragma omp parallel for schedule(static)
        for (k = 0; k < a->nnz; ++k) tmp[k] = (float)a->v[k];
        if ((rc = skk_malloc((void **)&d->vf, sizeof(float) * (size_t)(a->nnz ? a->nnz : 1))) ||
            (rc = skk_h2d(d->vf, tmp, sizeof(float) * (size_t)a->nnz))) { free(tmp); goto fail; }
        free(tmp);
    }
    free(tiles);
    d->k.rp = d->rp; d->k.ci = d->ci; d->k.vd = d->vd; d->k.vf = d->vf; d->k.tiles = d->tiles;
    return 0;
fail:
    free(tiles);
    dcsr_free(d);
    return skg_hip_fail("device CSR upload", rc);
}
