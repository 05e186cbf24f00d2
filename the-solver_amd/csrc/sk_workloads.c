/* Seeded synthetic workloads; see sk_workloads.h for the laws. */
#include "sk_workloads.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ PRNG */

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed)
{
    int k;
    for (k = 0; k < 4; ++k) r->s[k] = splitmix64(&seed);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t *r)   /* xoshiro256** */
{
    uint64_t *s = r->s;
    const uint64_t result = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform in [0, 1) from the top 53 bits */
static inline double rng_uniform(rng_t *r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

/* uniform integer in [0, n) by rejection (no modulo bias) */
static inline uint32_t rng_below(rng_t *r, uint32_t n)
{
    const uint64_t limit = UINT64_MAX - (UINT64_MAX % n);
    uint64_t v;
    do v = rng_next(r); while (v >= limit);
    return (uint32_t)(v % n);
}

/* N(0,1) by the Box-Muller cosine branch; u1 in (0,1] avoids log(0) */
static inline double rng_normal(rng_t *r)
{
    const double u1 = 1.0 - rng_uniform(r);
    const double u2 = rng_uniform(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

/* ----------------------------------------------------------------- model */

void skw_model_free(sk_model *m)
{
    if (!m) return;
    free(m->A.p); free(m->A.i); free(m->A.x);
    if (m->Q) { free(m->Q->p); free(m->Q->i); free(m->Q->x); free(m->Q); }
    free(m->c); free(m->rlow); free(m->rupp); free(m->clow); free(m->cupp);
    free(m->vartype);
    memset(m, 0, sizeof(*m));
    m->objsense = 1.0;
}

static int model_alloc(sk_model *m, int nrow, int ncol, long long nzmax)
{
    memset(m, 0, sizeof(*m));
    m->objsense = 1.0;
    m->nrow = nrow; m->ncol = ncol;
    m->A.nrow = nrow; m->A.ncol = ncol; m->A.nzmax = (int)nzmax;
    m->A.p = (int *)calloc((size_t)ncol + 1, sizeof(int));
    m->A.i = (int *)malloc((size_t)(nzmax > 0 ? nzmax : 1) * sizeof(int));
    m->A.x = (double *)malloc((size_t)(nzmax > 0 ? nzmax : 1) * sizeof(double));
    m->c = (double *)calloc((size_t)ncol + 1, sizeof(double));
    m->clow = (double *)calloc((size_t)ncol + 1, sizeof(double));
    m->cupp = (double *)calloc((size_t)ncol + 1, sizeof(double));
    m->rlow = (double *)calloc((size_t)nrow + 1, sizeof(double));
    m->rupp = (double *)calloc((size_t)nrow + 1, sizeof(double));
    m->vartype = (unsigned char *)calloc((size_t)ncol + 1, 1);
    if (!m->A.p || !m->A.i || !m->A.x || !m->c || !m->clow || !m->cupp ||
        !m->rlow || !m->rupp || !m->vartype) { skw_model_free(m); return -1; }
    return 0;
}

static int cmp_int(const void *a, const void *b)
{
    const int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* Draw `d` distinct rows in [0,nrow) sorted ascending into rows[]. */
static void distinct_rows(rng_t *r, int nrow, int d, int *rows)
{
    int k, again = 1;
    for (k = 0; k < d; ++k) rows[k] = (int)rng_below(r, (uint32_t)nrow);
    while (again) {
        again = 0;
        qsort(rows, (size_t)d, sizeof(int), cmp_int);
        for (k = 1; k < d; ++k)
            if (rows[k] == rows[k - 1]) { rows[k] = (int)rng_below(r, (uint32_t)nrow); again = 1; }
    }
}

static void csc_times(const sk_csc *a, const double *x, double *y)
{
    int j, p;
    memset(y, 0, (size_t)a->nrow * sizeof(double));
    for (j = 0; j < a->ncol; ++j)
        for (p = a->p[j]; p < a->p[j + 1]; ++p) y[a->i[p]] += a->x[p] * x[j];
}

/* --------------------------------------------------------- bounded LP */

int skw_bounded_lp(sk_model *m, int nrow, int ncol, int per_col, unsigned long long seed)
{
    rng_t r;
    int j, k, p = 0;
    double *x0, *ax;
    if (!m || nrow <= 0 || ncol <= 0 || per_col <= 0 || per_col > nrow) return -1;
    if (model_alloc(m, nrow, ncol, (long long)ncol * per_col)) return -1;
    rng_seed(&r, seed);
    x0 = (double *)malloc((size_t)ncol * sizeof(double));
    ax = (double *)malloc((size_t)nrow * sizeof(double));
    if (!x0 || !ax) { free(x0); free(ax); skw_model_free(m); return -1; }
    for (j = 0; j < ncol; ++j) {
        m->A.p[j] = p;
        distinct_rows(&r, nrow, per_col, m->A.i + p);
        for (k = 0; k < per_col; ++k) m->A.x[p + k] = rng_normal(&r);
        p += per_col;
    }
    m->A.p[ncol] = p;
    for (j = 0; j < ncol; ++j) x0[j] = rng_uniform(&r);
    csc_times(&m->A, x0, ax);
    for (k = 0; k < nrow; ++k) {
        m->rlow[k] = ax[k] - rng_uniform(&r);
        m->rupp[k] = ax[k] + rng_uniform(&r);
    }
    for (j = 0; j < ncol; ++j) {
        m->clow[j] = 0.0; m->cupp[j] = 10.0;
        m->c[j] = 2.0 * rng_uniform(&r) - 1.0;
    }
    free(x0); free(ax);
    return 0;
}

/* ------------------------------------------------------- power-law LP */

static double shifted_zipf_mean(double s, double q, int dmax)
{
    double num = 0.0, den = 0.0;
    int k;
    for (k = dmax; k >= 1; --k) {          /* small terms first */
        const double w = pow((double)k + q, -s);
        num += (double)k * w; den += w;
    }
    return num / den;
}

int skw_zipf_lp(sk_model *m, int nrow, int ncol, double s, double mean_degree,
                int max_degree, double eq_fraction, unsigned long long seed, double *q_out)
{
    rng_t r;
    double lo = -1.0 + 1e-12, hi = 0.0, q, *cdf, *x0, *ax;
    int *deg, j, k, it;
    long long nnz = 0, p = 0;
    if (!m || nrow <= 0 || ncol <= 0 || max_degree < 1 || max_degree > nrow) return -1;
    if (mean_degree <= 1.0 || mean_degree >= shifted_zipf_mean(s, 0.0, max_degree)) return -1;
    /* mean is increasing in q on (-1, 0]: bisect for the requested mean */
    for (it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (shifted_zipf_mean(s, mid, max_degree) < mean_degree) lo = mid; else hi = mid;
    }
    q = 0.5 * (lo + hi);
    if (q_out) *q_out = q;
    cdf = (double *)malloc((size_t)max_degree * sizeof(double));
    deg = (int *)malloc((size_t)ncol * sizeof(int));
    if (!cdf || !deg) { free(cdf); free(deg); return -1; }
    {
        double acc = 0.0;
        for (k = 0; k < max_degree; ++k) { acc += pow((double)(k + 1) + q, -s); cdf[k] = acc; }
        for (k = 0; k < max_degree; ++k) cdf[k] /= acc;
        cdf[max_degree - 1] = 1.0;
    }
    rng_seed(&r, seed);
    for (j = 0; j < ncol; ++j) {
        const double u = rng_uniform(&r);
        int a = 0, b = max_degree - 1;
        while (a < b) { const int c = (a + b) / 2; if (cdf[c] > u) b = c; else a = c + 1; }
        deg[j] = a + 1;
        nnz += deg[j];
    }
    free(cdf);
    if (nnz > 2147483647LL || model_alloc(m, nrow, ncol, nnz)) { free(deg); return -1; }
    for (j = 0; j < ncol; ++j) {
        m->A.p[j] = (int)p;
        distinct_rows(&r, nrow, deg[j], m->A.i + p);
        for (k = 0; k < deg[j]; ++k) {
            const double mag = pow(10.0, -3.0 + 6.0 * rng_uniform(&r));
            m->A.x[p + k] = (rng_next(&r) >> 63) ? -mag : mag;
        }
        p += deg[j];
    }
    m->A.p[ncol] = (int)p;
    free(deg);
    x0 = (double *)malloc((size_t)ncol * sizeof(double));
    ax = (double *)malloc((size_t)nrow * sizeof(double));
    if (!x0 || !ax) { free(x0); free(ax); skw_model_free(m); return -1; }
    for (j = 0; j < ncol; ++j) x0[j] = rng_uniform(&r);
    csc_times(&m->A, x0, ax);
    for (k = 0; k < nrow; ++k) {
        if (rng_uniform(&r) < eq_fraction) {
            m->rlow[k] = m->rupp[k] = ax[k];
        } else {
            m->rlow[k] = ax[k] - rng_uniform(&r);
            m->rupp[k] = ax[k] + rng_uniform(&r);
        }
    }
    for (j = 0; j < ncol; ++j) {
        m->clow[j] = 0.0; m->cupp[j] = 10.0;
        m->c[j] = 2.0 * rng_uniform(&r) - 1.0;
    }
    free(x0); free(ax);
    return 0;
}

/* ----------------------------------------------------------- sparse QP */

typedef struct { int col, row; double v; } triplet_t;

static int cmp_triplet(const void *a, const void *b)
{
    const triplet_t *x = (const triplet_t *)a, *y = (const triplet_t *)b;
    if (x->col != y->col) return (x->col > y->col) - (x->col < y->col);
    return (x->row > y->row) - (x->row < y->row);
}

int skw_sparse_qp(sk_model *m, int nrow, int ncol, int a_per_col, int b_rows,
                  int b_per_row, unsigned long long seed)
{
    rng_t r;
    int j, k, a, b;
    long long t = 0, ntrip, out;
    triplet_t *trip;
    int *cols;
    double *vals, *x0, *ax;
    sk_csc *Q;
    if (!m || nrow <= 0 || ncol <= 0 || a_per_col <= 0 || a_per_col > nrow ||
        b_rows < 0 || b_per_row <= 0 || b_per_row > ncol) return -1;
    if (skw_bounded_lp(m, nrow, ncol, a_per_col, seed)) return -1;
    /* the bounded-LP draws are reused for A; the QP-specific draws continue
       on a second stream so A is identical to the LP generator's */
    rng_seed(&r, seed ^ 0x5151515151515151ULL);
    ntrip = (long long)ncol + (long long)b_rows * b_per_row * b_per_row;
    trip = (triplet_t *)malloc((size_t)ntrip * sizeof(triplet_t));
    cols = (int *)malloc((size_t)b_per_row * sizeof(int));
    vals = (double *)malloc((size_t)b_per_row * sizeof(double));
    Q = (sk_csc *)calloc(1, sizeof(sk_csc));
    if (!trip || !cols || !vals || !Q) { free(trip); free(cols); free(vals); free(Q); skw_model_free(m); return -1; }
    for (j = 0; j < ncol; ++j) { trip[t].col = j; trip[t].row = j; trip[t].v = 0.1 + 9.9 * rng_uniform(&r); ++t; }
    for (k = 0; k < b_rows; ++k) {
        distinct_rows(&r, ncol, b_per_row, cols);
        for (a = 0; a < b_per_row; ++a) vals[a] = rng_normal(&r);
        for (a = 0; a < b_per_row; ++a)
            for (b = 0; b < b_per_row; ++b) {
                trip[t].col = cols[a]; trip[t].row = cols[b]; trip[t].v = vals[a] * vals[b]; ++t;
            }
    }
    qsort(trip, (size_t)t, sizeof(triplet_t), cmp_triplet);
    /* count distinct (col,row) */
    out = 0;
    for (k = 0; k < t; ++k) if (k == 0 || cmp_triplet(&trip[k], &trip[k - 1]) != 0) ++out;
    Q->nrow = Q->ncol = ncol; Q->nzmax = (int)out;
    Q->p = (int *)calloc((size_t)ncol + 1, sizeof(int));
    Q->i = (int *)malloc((size_t)out * sizeof(int));
    Q->x = (double *)malloc((size_t)out * sizeof(double));
    if (!Q->p || !Q->i || !Q->x) {
        free(Q->p); free(Q->i); free(Q->x); free(Q); free(trip); free(cols); free(vals);
        skw_model_free(m); return -1;
    }
    out = -1;
    for (k = 0; k < t; ++k) {
        if (k == 0 || cmp_triplet(&trip[k], &trip[k - 1]) != 0) {
            ++out; Q->i[out] = trip[k].row; Q->x[out] = trip[k].v; Q->p[trip[k].col + 1]++;
        } else {
            Q->x[out] += trip[k].v;
        }
    }
    for (j = 0; j < ncol; ++j) Q->p[j + 1] += Q->p[j];
    free(trip); free(cols); free(vals);
    m->Q = Q;
    /* QP data: c ~ N(0,1), bounds [-5,5], rows around A x0 with x0 ~ U[-1,1] */
    x0 = (double *)malloc((size_t)ncol * sizeof(double));
    ax = (double *)malloc((size_t)nrow * sizeof(double));
    if (!x0 || !ax) { free(x0); free(ax); skw_model_free(m); return -1; }
    for (j = 0; j < ncol; ++j) x0[j] = 2.0 * rng_uniform(&r) - 1.0;
    csc_times(&m->A, x0, ax);
    for (k = 0; k < nrow; ++k) {
        m->rlow[k] = ax[k] - rng_uniform(&r);
        m->rupp[k] = ax[k] + rng_uniform(&r);
    }
    for (j = 0; j < ncol; ++j) {
        m->clow[j] = -5.0; m->cupp[j] = 5.0;
        m->c[j] = rng_normal(&r);
    }
    free(x0); free(ax);
    return 0;
}

/* --------------------------------------------------------------- batch */

int skw_batch_sizes(int count, int m_lo, int m_hi, unsigned long long seed, int *m_out)
{
    rng_t r;
    int i;
    if (count < 0 || m_lo <= 0 || m_hi < m_lo || !m_out) return -1;
    rng_seed(&r, seed);
    for (i = 0; i < count; ++i) m_out[i] = m_lo + (int)rng_below(&r, (uint32_t)(m_hi - m_lo + 1));
    return 0;
}

/* ---------------------------------------------------------- parametric */

int skw_parametric_family(sk_model *m, int nrow, int ncol, int per_col, int dim,
                          unsigned long long seed, double *P, double *R,
                          double *c0, double *b0)
{
    rng_t r;
    int j, k;
    double *x0;
    if (skw_bounded_lp(m, nrow, ncol, per_col, seed)) return -1;
    rng_seed(&r, seed ^ 0xA5A5A5A5A5A5A5A5ULL);
    x0 = (double *)malloc((size_t)ncol * sizeof(double));
    if (!x0) { skw_model_free(m); return -1; }
    for (j = 0; j < ncol; ++j) x0[j] = rng_uniform(&r);
    csc_times(&m->A, x0, b0);
    free(x0);
    for (j = 0; j < ncol; ++j) c0[j] = 2.0 * rng_uniform(&r) - 1.0;
    for (k = 0; k < ncol * dim; ++k) P[k] = 0.1 * rng_normal(&r);
    for (k = 0; k < nrow * dim; ++k) R[k] = 0.1 * rng_normal(&r);
    for (j = 0; j < ncol; ++j) m->c[j] = c0[j];
    for (k = 0; k < nrow; ++k) { m->rlow[k] = b0[k] - 0.5; m->rupp[k] = b0[k] + 0.5; }
    return 0;
}

void skw_parametric_instance(int nrow, int ncol, int dim, const double *P, const double *R,
                             const double *c0, const double *b0, const double *theta,
                             double *c, double *rlow, double *rupp)
{
    int j, k, d;
    for (j = 0; j < ncol; ++j) {
        double acc = c0[j];
        for (d = 0; d < dim; ++d) acc += P[(size_t)j * dim + d] * theta[d];
        c[j] = acc;
    }
    for (k = 0; k < nrow; ++k) {
        double acc = 0.0, half;
        for (d = 0; d < dim; ++d) acc += R[(size_t)k * dim + d] * theta[d];
        half = 0.5 + 0.1 * fabs(acc);
        rlow[k] = b0[k] - half;
        rupp[k] = b0[k] + half;
    }
}

void skw_normal_matrix(int count, int dim, unsigned long long seed, double *out)
{
    rng_t r;
    long long k;
    rng_seed(&r, seed);
    for (k = 0; k < (long long)count * dim; ++k) out[k] = rng_normal(&r);
}
